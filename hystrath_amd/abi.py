"""ctypes mirror of include/hystrath_dsmc.h (the C ABI of the engine).

The structures here must match the header field for field; the
`test_abi_layout` test compares their sizes and offsets with what the
compiled library reports.
"""
import ctypes as C

HS_MAX_SPECIES = 8
HS_MAX_VIB_MODES = 3
HS_KB = 1.38064852e-23
HS_FILL_NOALLOC_MAX_PAIRS = 500000

HS_OK = 0
HS_E_INVALID = -1
HS_E_CAPACITY = -2
HS_E_HIP = -3
HS_E_OVERFLOW = -4
HS_E_UNSUPPORTED = -5

HS_COLL_NONE, HS_COLL_VHS, HS_COLL_VSS, HS_COLL_LB_VHS, HS_COLL_LB_VSS = range(5)
HS_PATCH_CYCLIC, HS_PATCH_SPECULAR, HS_PATCH_DIFFUSE, HS_PATCH_DELETION, HS_PATCH_EMPTY = range(5)

HS_FLAG_SAMPLE = 0x1
HS_FLAG_HEATFLUX = 0x2
HS_FLAG_COLL_SEP = 0x4
HS_FLAG_TIMING = 0x8
HS_FLAG_WALL = 0x10
HS_FLAG_FACE_FLUX = 0x20

HS_XMIN, HS_XMAX, HS_YMIN, HS_YMAX, HS_ZMIN, HS_ZMAX = range(6)

# sampling slots (enum hs_sample_slot / hs_species_slot)
HS_S_NCOLL, HS_S_COLLSEP, HS_S_GLOBAL_COUNT = 0, 1, 2
(HS_SS_N, HS_SS_MUX, HS_SS_MUY, HS_SS_MUZ, HS_SS_MCC, HS_SS_EROT, HS_SS_EVIB0) = range(7)
HS_SS_COUNT = HS_SS_EVIB0 + HS_MAX_VIB_MODES
HS_HF_COUNT = 13
# wall-face sums, [species*HS_WS_COUNT + slot][wall face]
(HS_WS_RHON, HS_WS_KE, HS_WS_MOMX, HS_WS_MOMY, HS_WS_MOMZ, HS_WS_EROT, HS_WS_EVIB0) = range(7)
HS_WS_Q = HS_WS_EVIB0 + HS_MAX_VIB_MODES
HS_WS_FDX, HS_WS_FDY, HS_WS_FDZ = HS_WS_Q + 1, HS_WS_Q + 2, HS_WS_Q + 3
HS_WS_COUNT = HS_WS_FDZ + 1
HS_WALL_CYLINDER = 6  # side value of the cylinder's wall faces


def species_slot(s, k):
    """Row of the sample array holding quantity k of species s."""
    return HS_S_GLOBAL_COUNT + s * HS_SS_COUNT + k

# names used by the reference dictionaries -> enums
COLLISION_MODELS = {
    "NoBinaryCollision": HS_COLL_NONE,
    "VariableHardSphere": HS_COLL_VHS,
    "VariableSoftSphere": HS_COLL_VSS,
    "LarsenBorgnakkeVariableHardSphere": HS_COLL_LB_VHS,
    "LarsenBorgnakkeVariableSoftSphere": HS_COLL_LB_VSS,
}
PATCH_MODELS = {
    "cyclic": HS_PATCH_CYCLIC,
    "dsmcSpecularWallPatch": HS_PATCH_SPECULAR,
    "dsmcDiffuseWallPatch": HS_PATCH_DIFFUSE,
    "dsmcDeletionPatch": HS_PATCH_DELETION,
    "empty": HS_PATCH_EMPTY,
}


class hs_species(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("diameter", C.c_double),
        ("omega", C.c_double),
        ("alpha", C.c_double),
        ("rot_dof", C.c_double),
        ("n_vib_modes", C.c_int32),
        ("charge", C.c_int32),
        ("theta_v", C.c_double * HS_MAX_VIB_MODES),
        ("zref", C.c_double * HS_MAX_VIB_MODES),
        ("tref_zv", C.c_double * HS_MAX_VIB_MODES),
        ("theta_d", C.c_double),
    ]


class hs_patch(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("inflow", C.c_int32),
        ("temperature", C.c_double),
        ("velocity", C.c_double * 3),
    ]


HS_REACT_DISSOCIATION_QK = 1
HS_REACT_EXCHANGE_QK = 2
HS_REACT_DISSOCIATION_EXCHANGE_QK = 3
HS_MAX_REACTIONS = 32

# reactionModel names of chemReactDict (dsmcReaction.C:106-135)
REACTION_MODELS = {
    "dissociationQK": HS_REACT_DISSOCIATION_QK,
    "exchangeQK": HS_REACT_EXCHANGE_QK,
    "dissociationExchangeQK": HS_REACT_DISSOCIATION_EXCHANGE_QK,
}


class hs_reaction(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("reactants", C.c_int32 * 2),
        ("diss_products", (C.c_int32 * 2) * 2),
        ("exch_products", C.c_int32 * 2),
        ("heat_exchange", C.c_double),
        ("a_coeff", C.c_double),
        ("b_coeff", C.c_double),
        ("allow_splitting", C.c_int32),
    ]


class hs_config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("origin", C.c_double * 3),
        ("delta", C.c_double * 3),
        ("cell_volume", C.POINTER(C.c_double)),
        ("patch", hs_patch * 6),
        ("cyl_enabled", C.c_int32),
        ("cyl_center", C.c_double * 2),
        ("cyl_radius", C.c_double),
        ("cyl_patch", hs_patch),
        ("n_species", C.c_int32),
        ("species", hs_species * HS_MAX_SPECIES),
        ("collision_model", C.c_int32),
        ("tref", C.c_double),
        ("z_rot", C.c_double),
        ("z_vib_fixed", C.c_double),
        ("z_el", C.c_double),
        ("inv_zv_form", C.c_int32),
        ("inflow_n", C.c_double * HS_MAX_SPECIES),
        ("inflow_t", C.c_double),
        ("inflow_trot", C.c_double),
        ("inflow_tvib", C.c_double),
        ("inflow_u", C.c_double * 3),
        ("fn", C.c_double),
        ("dt", C.c_double),
        ("seed", C.c_uint64),
        ("capacity", C.c_int64),
        ("flags", C.c_int32),
        ("reactions", C.POINTER(hs_reaction)),
        ("n_reactions", C.c_int32),
    ]


class hs_particles(C.Structure):
    _fields_ = [
        ("px", C.POINTER(C.c_double)), ("py", C.POINTER(C.c_double)), ("pz", C.POINTER(C.c_double)),
        ("ux", C.POINTER(C.c_float)), ("uy", C.POINTER(C.c_float)), ("uz", C.POINTER(C.c_float)),
        ("cell", C.POINTER(C.c_int32)),
        ("erot", C.POINTER(C.c_float)),
        ("vib", C.POINTER(C.c_uint16)),
        ("type", C.POINTER(C.c_uint8)),
    ]


def ptr(arr, ctype):
    """ctypes pointer to a numpy array's data (None for None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


def sample_slots(cfg):
    n = HS_S_GLOBAL_COUNT + cfg.n_species * HS_SS_COUNT
    if cfg.flags & HS_FLAG_HEATFLUX:
        n += cfg.n_species * HS_HF_COUNT
    return n
