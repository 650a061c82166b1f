"""MI355X-native space-charge PIC step (OPAL-T drop-in) -- HIP kernels behind a C ABI."""
from .opalsc import (BC_OPEN, BC_PERIODIC, Batch, COUPLING, CLIGHT, EXPORTS, FP64, MIXED,  # noqa: F401
                     GREENS_INTEGRATED,
                     GREENS_STANDARD, GRID_EFIELD, GRID_PHI, GRID_RHO, M_P_EV, FFTPoissonSolver,
                     OpalscError, SpaceChargeContext, StatRow, fromfile_read, host_atan_cr, host_green_octant,
                     host_igf_term, host_log_cr, lib,
                     stat_write, tune_deposit_get, tune_deposit_set, tune_fft_candidates, tune_fft_get,
                     tune_fft_set, tune_sort_get, tune_sort_set)
