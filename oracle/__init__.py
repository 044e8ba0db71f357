"""CPU oracle for the reinit path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's cpu_baseline leg may use this
package; the product (fluidtissue_amd / libftreinit.so) never does.
"""
from .oracle import *  # noqa: F401,F403
