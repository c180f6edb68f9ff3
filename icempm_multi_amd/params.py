"""SimParams: ctypes mirror of ``icempm_simparams`` (include/icempm.h).

That struct is itself a field-for-field copy of the reference's POD
``icy::SimParams`` (simulation/parameters_sim.h:37-85), which the reference
copies by ``sizeof`` to device constant memory (gpu_partition.cu:373).  The
field order and C types below are therefore fixed by that struct.
"""
import ctypes as C

_FIELDS = [
    ("n_indenter_subdivisions", C.c_int),
    ("tpb_P2G", C.c_int), ("tpb_Upd", C.c_int), ("tpb_G2P", C.c_int),
    ("nPtsTotal", C.c_int),
    ("GridXTotal", C.c_int), ("GridY", C.c_int), ("GridTotal", C.c_int),
    ("GridXDimension", C.c_double),
    ("InitialTimeStep", C.c_double), ("SimulationEndTime", C.c_double),
    ("UpdateEveryNthStep", C.c_int),
    ("SimulationStep", C.c_int),
    ("SimulationTime", C.c_double),
    ("Gravity", C.c_double), ("Density", C.c_double), ("WaterDensity", C.c_double),
    ("PoissonsRatio", C.c_double), ("YoungsModulus", C.c_double),
    ("lambda_", C.c_double), ("mu", C.c_double), ("kappa", C.c_double),
    ("IceCompressiveStrength", C.c_double), ("IceTensileStrength", C.c_double),
    ("IceShearStrength", C.c_double), ("IceTensileStrength2", C.c_double),
    ("NACC_beta", C.c_double), ("NACC_M", C.c_double), ("NACC_Msq", C.c_double),
    ("DP_tan_phi", C.c_double), ("DP_threshold_p", C.c_double),
    ("IndDiameter", C.c_double), ("IndRSq", C.c_double), ("IndVelocity", C.c_double),
    ("IndDepth", C.c_double),
    ("cellsize", C.c_double), ("cellsize_inv", C.c_double), ("Dp_inv", C.c_double),
    ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
    ("nxmin", C.c_int), ("nxmax", C.c_int), ("nymin", C.c_int), ("nymax", C.c_int),
    ("ParticleVolume", C.c_double), ("ParticleMass", C.c_double),
    ("WaterParticleMass", C.c_double), ("ParticleViewSize", C.c_double),
    ("indenter_x", C.c_double), ("indenter_x_initial", C.c_double),
    ("indenter_y", C.c_double), ("indenter_y_initial", C.c_double),
    ("Volume", C.c_double),
    ("SetupType", C.c_int),
    ("GrainVariability", C.c_double),
    ("GridHaloSize", C.c_int),
    ("ExtraSpaceForIncomingPoints", C.c_double),
    ("PointsTransferBufferFraction", C.c_double),
    ("RebalanceThresholdFreeSpaceRemaining", C.c_double),
    ("RebalanceThresholdDisabledPercentage", C.c_double),
    ("PointTransferFrequency", C.c_int),
    ("nPartitions", C.c_int),
    ("VectorCapacity_transfer", C.c_int),
    ("dt_vol_Dpinv", C.c_double), ("dt_Gravity", C.c_double),
    ("vmax", C.c_double), ("vmax_squared", C.c_double),
    ("gbOffset", C.c_int),
]


class SimParams(C.Structure):
    """``icempm_simparams``; ``lambda_`` stands for the C member ``lambda``."""
    _fields_ = _FIELDS

    def copy(self):
        other = SimParams()
        C.memmove(C.byref(other), C.byref(self), C.sizeof(SimParams))
        return other

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in _FIELDS}


FIELD_NAMES = [name for name, _ in _FIELDS]

# HostSideSOA line indices (parameters_sim.h:27-35)
IDX_UTILITY_DATA = 0
IDX_P = 1
IDX_Q = 2
IDX_JP_INV = 3
POSX = 4
VELX = 6
FE00 = 8
BP00 = 12
N_PTS_ARRAYS = 16

STATUS_CRUSHED = 0x10000
STATUS_DISABLED = 0x20000
STATUS_LIQUID = 0x40000

PRECISION_FP64 = 0
PRECISION_MIXED = 1

STEP_GRAPH = 0x1
STEP_RESORT = 0x2
STEP_TIMINGS = 0x4
STEP_DIRECT_P2G = 0x8
STEP_UNFUSED = 0x10
STEP_SAMPLED = 0x20
STEP_SPAN = 0x40

# status codes of every icempm_* call (ICEMPM_OK, ICEMPM_ERR_* of include/icempm.h)
OK = 0
ERR_ARGUMENT = -1
ERR_HIP = -2
ERR_CAPACITY = -3
ERR_STATE = -4
ERR_NOMEM = -5
ERR_DEVICE = -6
