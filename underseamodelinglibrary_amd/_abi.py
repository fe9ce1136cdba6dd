"""ctypes mirror of include/wq3d.h (the C ABI of libwq3d.so).

Field order and types must match the header exactly; tests/test_abi.py checks
the struct sizes against offsets the C compiler reports.
"""
import ctypes as C

WQ3D_OK = 0
WQ3D_ERR_ARGUMENT = 1
WQ3D_ERR_UNSUPPORTED = 2
WQ3D_ERR_DEVICE = 3
WQ3D_ERR_CAPACITY = 4
WQ3D_ERR_STATE = 5

SEQ_LINEAR, SEQ_LOG, SEQ_DATA = 0, 1, 2
INTERP_NEAREST, INTERP_LINEAR, INTERP_PCHIP = 0, 1, 2
PROFILE_LINEAR, PROFILE_MUNK, PROFILE_GRID = 0, 1, 2
ATTEN_THORP, ATTEN_CONSTANT = 0, 1
BOUNDARY_FLAT, BOUNDARY_GRID = 0, 1
LOSS_CONSTANT, LOSS_RAYLEIGH = 0, 1
SPREADING_CLASSIC_RAY, SPREADING_HYBRID_GAUSSIAN = 0, 1
PAST, PREV, CURR, NEXT = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_uint8)


class Seq(C.Structure):
    _fields_ = [("size", C.c_int64), ("value", _dp), ("increment", _dp),
                ("search", C.c_int32), ("pad_", C.c_int32)]


class Grid(C.Structure):
    _fields_ = [("ndims", C.c_int32), ("interp", C.c_int32 * 3), ("edge_limit", C.c_int32 * 3),
                ("pad_", C.c_int32), ("axis", Seq * 3), ("data", _dp)]


class ProfileDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flat_earth", C.c_int32), ("p", C.c_double * 4),
                ("grid", Grid), ("atten_kind", C.c_int32), ("pad_", C.c_int32),
                ("atten_coeff", C.c_double)]


class BoundaryDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("loss_kind", C.c_int32), ("rho", C.c_double),
                ("normal_rho", C.c_double), ("loss", C.c_double * 5), ("grid", Grid)]


class VolumeDesc(C.Structure):
    _fields_ = [("rho", C.c_double), ("thickness", C.c_double)]


class EnvDesc(C.Structure):
    _fields_ = [("profile", ProfileDesc), ("surface", BoundaryDesc), ("bottom", BoundaryDesc),
                ("num_volume", C.c_int32), ("pad_", C.c_int32),
                ("volume", C.POINTER(VolumeDesc))]


class BatchShape(C.Structure):
    _fields_ = [("ray_threads", C.c_int32), ("build_blocks", C.c_int32), ("use_graph", C.c_int32),
                ("pad_", C.c_int32)]


class Thresholds(C.Structure):
    _fields_ = [("intensity", C.c_double), ("eigenverb", C.c_double),
                ("max_bottom", C.c_int32), ("max_surface", C.c_int32),
                ("max_caustic", C.c_int32), ("max_upper", C.c_int32),
                ("max_lower", C.c_int32), ("pad_", C.c_int32)]


class FanDesc(C.Structure):
    _fields_ = [("freq", Seq), ("de", Seq), ("az", Seq),
                ("src_rho", C.c_double), ("src_theta", C.c_double), ("src_phi", C.c_double),
                ("time_step", C.c_double), ("earth_radius", C.c_double),
                ("num_t1", C.c_int64), ("num_t2", C.c_int64),
                ("target_rho", _dp), ("target_theta", _dp), ("target_phi", _dp),
                ("spreading", C.c_int32), ("want_eigenverbs", C.c_int32),
                ("want_reflections", C.c_int32), ("pad_", C.c_int32),
                ("thresholds", Thresholds)]


class EigenrayRec(C.Structure):
    _fields_ = [("step", C.c_int64), ("t1", C.c_int32), ("t2", C.c_int32),
                ("de", C.c_int32), ("az", C.c_int32),
                ("travel_time", C.c_double), ("source_de", C.c_double),
                ("source_az", C.c_double), ("target_de", C.c_double),
                ("target_az", C.c_double),
                ("surface", C.c_int32), ("bottom", C.c_int32), ("caustic", C.c_int32),
                ("upper", C.c_int32), ("lower", C.c_int32), ("pad_", C.c_int32)]


class EigenrayTotal(C.Structure):
    _fields_ = [("travel_time", C.c_double), ("source_de", C.c_double),
                ("source_az", C.c_double), ("target_de", C.c_double),
                ("target_az", C.c_double), ("initial_time", C.c_double),
                ("surface", C.c_int32), ("bottom", C.c_int32), ("caustic", C.c_int32),
                ("upper", C.c_int32), ("lower", C.c_int32), ("count", C.c_int32)]


class EigenverbRec(C.Structure):
    _fields_ = [("step", C.c_int64), ("de", C.c_int32), ("az", C.c_int32),
                ("seq", C.c_int32), ("type", C.c_int32),
                ("travel_time", C.c_double), ("length", C.c_double), ("width", C.c_double),
                ("direction", C.c_double), ("grazing", C.c_double),
                ("sound_speed", C.c_double),
                ("pos_rho", C.c_double), ("pos_theta", C.c_double), ("pos_phi", C.c_double),
                ("source_de", C.c_double), ("source_az", C.c_double),
                ("surface", C.c_int32), ("bottom", C.c_int32), ("caustic", C.c_int32),
                ("upper", C.c_int32), ("lower", C.c_int32), ("pad_", C.c_int32)]


class ReflectionRec(C.Structure):
    _fields_ = [("step", C.c_int64), ("de", C.c_int32), ("az", C.c_int32),
                ("seq", C.c_int32), ("type", C.c_int32),
                ("time", C.c_double), ("dt", C.c_double), ("grazing", C.c_double),
                ("speed", C.c_double),
                ("pos_rho", C.c_double), ("pos_theta", C.c_double), ("pos_phi", C.c_double),
                ("ndir_rho", C.c_double), ("ndir_theta", C.c_double), ("ndir_phi", C.c_double)]


FRONT_FIELDS = ["pos_rho", "pos_theta", "pos_phi", "pgrad_rho", "pgrad_theta", "pgrad_phi",
                "ndir_rho", "ndir_theta", "ndir_phi", "dgrad_rho", "dgrad_theta", "dgrad_phi",
                "sound_speed", "sgrad_rho", "sgrad_theta", "sgrad_phi", "distance",
                "path_length", "attenuation", "phase"]
FRONT_INT_FIELDS = ["surface", "bottom", "caustic", "upper", "lower"]


class FrontView(C.Structure):
    _fields_ = ([(n, _dp) for n in FRONT_FIELDS] + [(n, _ip) for n in FRONT_INT_FIELDS]
                + [("on_edge", _bp)])


# numpy dtypes that match the record structs, for zero-copy conversion
def record_dtype(struct):
    import numpy as np
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        off = getattr(struct, name).offset
        fmt = {C.c_int64: "<i8", C.c_int32: "<i4", C.c_double: "<f8"}[ctype]
        names.append(name)
        formats.append(fmt)
        offsets.append(off)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": C.sizeof(struct)})
