"""ctypes binding of libmra_hip.so (the C ABI declared in include/mra.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MRA_LIBRARY selects another build of the same library (same C ABI)
LIB_PATH = os.environ.get("MRA_LIBRARY") or os.path.join(HERE, "_build", "libmra_hip.so")

MRA_OK, MRA_E_ARG, MRA_E_HIP, MRA_E_NOT_PD, MRA_E_STATE, MRA_E_NOMEM, MRA_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
METRIC = {"chordal": 0, "great_circle": 1, "planar": 2}
METRIC_PLANAR = 2
FIELD_CONSTANT, FIELD_WENDLAND, FIELD_LINEAR = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
T_SETUP, T_FINEST_CROSS, T_FINEST_LOCAL, T_PRED_PRE, T_PRED_POST, T_REDUCE, T_GRAPH, T_PRIOR0, T_POST0, T_COUNT = \
    0, 1, 2, 3, 4, 5, 6, 8, 32, 56

u64p = C.POINTER(C.c_uint64)
dp = C.POINTER(C.c_double)

# Every symbol include/mra.h declares.
EXPORTED = ["mra_create", "mra_destroy", "mra_last_error", "mra_set_covariance", "mra_set_params",
            "mra_set_precision", "mra_likelihood", "mra_predict", "mra_likelihood_levels", "mra_sample_prior",
            "mra_set_observations", "mra_use_graph", "mra_last_timings", "mra_cross_covariance",
            "mra_variance_covariance", "mra_cov_tile_device", "mra_partition_program_sizes",
            "mra_partition_program", "mra_version", "mra_selftest", "mra_set_option", "mra_get_option", "mra_selftest_f32", "mra_selftest_blk",
            "mra_bobyqa", "mra_fit_linear_trend", "mra_variogram_features", "mra_mlp_create", "mra_mlp_destroy",
            "mra_mlp_last_error", "mra_mlp_param_count", "mra_mlp_set_params", "mra_mlp_get_params",
            "mra_mlp_forward", "mra_mlp_loss_grad", "mra_mlp_train_step", "mra_batch_create", "mra_batch_destroy",
            "mra_batch_set", "mra_batch_get", "mra_batch_likelihood", "mra_batch_autotune", "mra_set_tuning",
            "mra_get_tuning", "mra_autotune", "mra_cov_tile_timed", "mra_device_alloc", "mra_device_free",
            "mra_device_copy", "mra_batch_last_error", "mra_loglikelihood", "mra_cross_covariance_xyz",
            "mra_variance_covariance_xyz", "mra_fields", "mra_io_last_error", "mra_user_params_init",
            "mra_read_user_parameters", "mra_check_user_parameters", "mra_data_load", "mra_data_destroy",
            "mra_data_get", "mra_partition_build", "mra_partition_destroy", "mra_partition_tree",
            "mra_partition_indices", "mra_partition_level_counts", "mra_partition_dump_structure",
            "mra_default_levels", "mra_cov_model_load", "mra_cov_model_destroy", "mra_cov_model_desc",
            "mra_write_prediction_result", "mra_partition_program_gpu", "mra_partition_gpu_last_error",
            "mra_cross_covariance_noalloc", "mra_cross_covariance_noalloc_limit", "mra_fit_linear_trend_npt",
            "mra_bobyqa_h0"]
LEVELS_DEFAULT = -99
OPT_FACTORED_POSTERIOR = 1
OPT_GRAM_KERNEL = 2
OPT_WAVE_LOCAL16 = 3
OPT_CROSS16 = 4
OPT_CONCURRENT = 5
OPT_SUM_MESSAGES = 6
OPT_XCD_MAP = 7
OPT_MIXED_FP64_ABOVE = 8
OPT_PERSISTENT = 9
OPT_PRIOR_WAVE = 10
OPT_FINEST_WAVE = 11
OPT_LOCAL_LL = 12


class TreeDesc(C.Structure):
    _fields_ = [("num_levels", C.c_int32), ("num_partitions", C.c_int32), ("num_knots", C.c_int32),
                ("various_j", C.c_int32), ("n_coarse", C.c_uint64), ("n_finest", C.c_uint64),
                ("n_obs", C.c_uint64), ("n_pred", C.c_uint64), ("knot_lon", dp), ("knot_lat", dp),
                ("finest_parent", u64p), ("obs_offsets", u64p), ("obs_lon", dp), ("obs_lat", dp),
                ("obs_val", dp), ("pred_offsets", u64p), ("pred_lon", dp), ("pred_lat", dp)]


class FieldDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("n_basis", C.c_uint64), ("intercept", C.c_double),
                ("basis_lon", dp), ("basis_lat", dp), ("basis_weight", dp), ("slope_x", C.c_double),
                ("slope_y", C.c_double)]


class CovDesc(C.Structure):
    _fields_ = [("nonstationary", C.c_int32), ("metric", C.c_int32), ("wendland_len", C.c_double),
                ("sigmasq", FieldDesc), ("beta", FieldDesc), ("tausq", FieldDesc)]


_PATHS = ("data_file_name", "prediction_location_file", "prediction_results_file_name", "sigmasq_file_name",
          "beta_file_name", "tausq_file_name", "knots_file_name", "parent_region_file_name", "tmp_directory")
_FLAGS = ("elimination_duplicates_flag", "regression_flag", "nonstationary_flag", "print_detail_flag",
          "dump_prediction_results_flag", "save_to_disk_flag", "dynamic_schedule_flag", "chordal_distance_flag",
          "provided_knots_flag", "obs_as_knots_flag", "various_j_finest_level_flag")
_INTS = ("num_partitions_j", "num_knots_r", "num_levels_m", "num_knots_finest", "max_iterations")
_DBLS = ("offset", "wendland_len", "sigmasq", "beta", "tausq", "sigmasq_scale", "beta_scale", "tausq_scale",
         "sigmasq_initial_guess", "sigmasq_lower_bound", "sigmasq_upper_bound", "beta_initial_guess",
         "beta_lower_bound", "beta_upper_bound", "tausq_initial_guess", "tausq_lower_bound", "tausq_upper_bound")


class UserParams(C.Structure):
    """mra_user_params: the reference's run configuration (user_parameters keys)."""
    _fields_ = ([(k, C.c_char * 1024) for k in _PATHS] + [("calculation_mode", C.c_char * 64),
                                                           ("prediction_location_mode", C.c_char),
                                                           ("pad_", C.c_char * 3)] +
                [(k, C.c_int32) for k in _FLAGS + _INTS] + [(k, C.c_double) for k in _DBLS])

    def to_dict(self):
        d = {k: getattr(self, k).decode() for k in _PATHS + ("calculation_mode",)}
        d["prediction_location_mode"] = self.prediction_location_mode.decode()
        d.update({k: getattr(self, k) for k in _FLAGS + _INTS + _DBLS})
        return d

    @staticmethod
    def from_dict(d):
        p = UserParams()
        lib().mra_user_params_init(C.byref(p))
        for k, v in d.items():
            if k in _PATHS or k == "calculation_mode":
                setattr(p, k, str(v).encode())
            elif k == "prediction_location_mode":
                p.prediction_location_mode = str(v).encode()[:1]
            else:
                setattr(p, k, v)
        return p


class DataView(C.Structure):
    _fields_ = [("n_locations", C.c_uint64), ("n_obs_raw", C.c_uint64), ("n_pred_raw", C.c_uint64),
                ("n_obs", C.c_uint64), ("n_pred", C.c_uint64), ("obs_lon", dp), ("obs_lat", dp), ("obs_val", dp),
                ("obs_region", u64p), ("pred_lon", dp), ("pred_lat", dp), ("pred_region", u64p),
                ("bounds", C.c_double * 4), ("coefficients", C.c_double * 3), ("regression", C.c_int32),
                ("predicting", C.c_int32)]


class MraError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"MRA error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__; "
                               f"__graft_entry__.build()'` first (no CPU fallback exists)")
        _lib = C.CDLL(LIB_PATH)
        _lib.mra_last_error.restype = C.c_char_p
        _lib.mra_last_error.argtypes = [C.c_void_p]
        _lib.mra_version.restype = C.c_char_p
        _lib.mra_create.argtypes = [C.POINTER(TreeDesc), C.POINTER(C.c_void_p)]
        _lib.mra_destroy.argtypes = [C.c_void_p]
        _lib.mra_set_covariance.argtypes = [C.c_void_p, C.POINTER(CovDesc)]
        _lib.mra_set_params.argtypes = [C.c_void_p] + [C.c_double] * 6
        _lib.mra_set_precision.argtypes = [C.c_void_p, C.c_int]
        _lib.mra_likelihood.argtypes = [C.c_void_p, dp]
        _lib.mra_likelihood_levels.argtypes = [C.c_void_p, dp]
        _lib.mra_predict.argtypes = [C.c_void_p, dp, dp, dp]
        _lib.mra_sample_prior.argtypes = [C.c_void_p, C.c_uint64, dp, C.c_int]
        _lib.mra_set_observations.argtypes = [C.c_void_p, dp]
        _lib.mra_use_graph.argtypes = [C.c_void_p, C.c_int]
        _lib.mra_last_timings.argtypes = [C.c_void_p, dp, C.c_int]
        _lib.mra_cross_covariance.argtypes = [C.POINTER(CovDesc)] + [C.c_double] * 6 + [dp, dp, C.c_uint64, dp, dp,
                                                                                       C.c_uint64, dp]
        _lib.mra_variance_covariance.argtypes = [C.POINTER(CovDesc)] + [C.c_double] * 6 + [dp, dp, C.c_uint64, dp]
        _lib.mra_cov_tile_device.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                             C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        _lib.mra_cov_tile_timed.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                            C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, dp]
        _lib.mra_device_alloc.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
        _lib.mra_device_free.argtypes = [C.c_void_p]
        _lib.mra_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
        _lib.mra_partition_program_sizes.argtypes = [dp, dp, C.c_uint64, dp, dp, C.c_uint64, C.c_int, C.c_int,
                                                     C.c_int, C.c_double, dp, u64p]
        _lib.mra_partition_program.argtypes = [dp, dp, dp, C.c_uint64, dp, dp, C.c_uint64, C.c_int, C.c_int,
                                               C.c_int, C.c_double, dp, dp, dp, u64p, u64p, dp, dp, dp, u64p, u64p,
                                               dp, dp, u64p]
        _lib.mra_set_option.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.mra_get_option.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        fp = C.POINTER(C.c_float)
        _lib.mra_selftest_f32.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, fp]
        _lib.mra_selftest.argtypes = [C.c_int, C.c_int, C.c_int] + [dp] * 8 + [C.POINTER(C.c_int)]
        _lib.mra_selftest_blk.argtypes = [C.c_int, dp, dp, dp, C.POINTER(C.c_int)]
        _lib.mra_bobyqa.argtypes = [C.c_int, C.c_int, OBJECTIVE, C.c_void_p, dp, dp, dp, C.c_double, C.c_double,
                                    C.c_int, dp, C.POINTER(C.c_int)]
        _lib.mra_fit_linear_trend.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_int, dp,
                                              C.POINTER(C.c_int)]
        _lib.mra_fit_linear_trend_npt.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                                                  C.c_int, dp, dp, C.POINTER(C.c_int)]
        _lib.mra_bobyqa_h0.argtypes = [C.c_int, C.c_int, OBJECTIVE, C.c_void_p, dp, dp, dp, C.c_double, C.c_double,
                                       C.c_int, dp, dp, C.POINTER(C.c_int)]
        _lib.mra_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        _lib.mra_get_tuning.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3
        _lib.mra_autotune.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3 + [dp]
        _lib.mra_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
        _lib.mra_batch_destroy.argtypes = [C.c_void_p]
        _lib.mra_batch_destroy.restype = None
        _lib.mra_batch_set.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.mra_batch_get.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _lib.mra_batch_likelihood.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int)]
        _lib.mra_batch_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), dp]
        _lib.mra_batch_last_error.argtypes = [C.c_void_p]
        _lib.mra_batch_last_error.restype = C.c_char_p
        fp_ = C.POINTER(C.c_float)
        _lib.mra_variogram_features.argtypes = [C.c_void_p, fp_]
        _lib.mra_mlp_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
        _lib.mra_mlp_destroy.argtypes = [C.c_void_p]
        _lib.mra_mlp_destroy.restype = None
        _lib.mra_mlp_last_error.argtypes = [C.c_void_p]
        _lib.mra_mlp_last_error.restype = C.c_char_p
        _lib.mra_mlp_set_params.argtypes = [C.c_void_p, fp_]
        _lib.mra_mlp_get_params.argtypes = [C.c_void_p, fp_]
        _lib.mra_mlp_forward.argtypes = [C.c_void_p, fp_, C.c_int, fp_, C.c_int]
        _lib.mra_mlp_loss_grad.argtypes = [C.c_void_p, fp_, fp_, C.c_int, fp_, fp_]
        _lib.mra_mlp_train_step.argtypes = [C.c_void_p, fp_, fp_, C.c_int, C.c_float, fp_]
        _lib.mra_loglikelihood.argtypes = [C.c_void_p, C.c_uint64, dp, dp]
        _lib.mra_cross_covariance_xyz.argtypes = [C.POINTER(CovDesc)] + [C.c_double] * 6 + [dp] * 5 + [C.c_uint64] + \
            [dp] * 5 + [C.c_uint64, dp]
        _lib.mra_variance_covariance_xyz.argtypes = [C.POINTER(CovDesc)] + [C.c_double] * 6 + [dp] * 5 + \
            [C.c_uint64, dp]
        _lib.mra_fields.argtypes = [C.POINTER(CovDesc), dp, dp, dp, dp, dp, C.c_uint64, dp, dp, dp]
        _lib.mra_io_last_error.restype = C.c_char_p
        _lib.mra_io_last_error.argtypes = []
        _lib.mra_user_params_init.argtypes = [C.POINTER(UserParams)]
        _lib.mra_user_params_init.restype = None
        _lib.mra_read_user_parameters.argtypes = [C.c_char_p, C.POINTER(UserParams)]
        _lib.mra_check_user_parameters.argtypes = [C.POINTER(UserParams)]
        _lib.mra_data_load.argtypes = [C.POINTER(UserParams), C.POINTER(C.c_void_p)]
        _lib.mra_data_destroy.argtypes = [C.c_void_p]
        _lib.mra_data_destroy.restype = None
        _lib.mra_data_get.argtypes = [C.c_void_p, C.POINTER(DataView)]
        _lib.mra_partition_build.argtypes = [C.POINTER(UserParams), C.c_void_p, C.POINTER(C.c_void_p)]
        _lib.mra_partition_destroy.argtypes = [C.c_void_p]
        _lib.mra_partition_destroy.restype = None
        _lib.mra_partition_tree.argtypes = [C.c_void_p, C.POINTER(TreeDesc)]
        _lib.mra_partition_indices.argtypes = [C.c_void_p, C.POINTER(u64p), C.POINTER(u64p)]
        _lib.mra_partition_level_counts.argtypes = [C.c_void_p, u64p]
        _lib.mra_partition_dump_structure.argtypes = [C.c_void_p, C.c_char_p]
        _lib.mra_default_levels.argtypes = [C.c_uint64, C.c_int, C.c_int]
        _lib.mra_cov_model_load.argtypes = [C.POINTER(UserParams), C.POINTER(C.c_void_p)]
        _lib.mra_cov_model_destroy.argtypes = [C.c_void_p]
        _lib.mra_cov_model_destroy.restype = None
        _lib.mra_cov_model_desc.argtypes = [C.c_void_p, C.POINTER(CovDesc)]
        _lib.mra_write_prediction_result.argtypes = [C.c_char_p, C.c_uint64, dp, dp, dp, dp, dp]
        _lib.mra_partition_program_gpu.argtypes = [dp, dp, dp, C.c_uint64, dp, dp, C.c_uint64, C.c_int, C.c_int,
                                                   C.c_int, C.c_double, dp, dp, dp, u64p, u64p, dp, dp, dp, u64p,
                                                   u64p, dp, dp, u64p, u64p]
        _lib.mra_partition_gpu_last_error.restype = C.c_char_p
        _lib.mra_cross_covariance_noalloc.argtypes = _lib.mra_cross_covariance.argtypes
        _lib.mra_cross_covariance_noalloc_limit.argtypes = []
        _lib.mra_cross_covariance_noalloc_limit.restype = C.c_uint64
        _lib.mra_partition_gpu_last_error.argtypes = []
    return _lib


def check_io(rc):
    if rc != MRA_OK:
        raise MraError(rc, lib().mra_io_last_error().decode())


def read_user_parameters(path, base=None):
    """mra_read_user_parameters: the `user_parameters` file as a dict (fields
    the file does not name keep `base`'s values, else zero)."""
    p = UserParams.from_dict(base or {})
    check_io(lib().mra_read_user_parameters(str(path).encode(), C.byref(p)))
    return p.to_dict()


def check_user_parameters(params):
    check_io(lib().mra_check_user_parameters(C.byref(UserParams.from_dict(params))))


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class DataSet:
    """mra_data: Data::load_data of a parameter dict (files, NaN split,
    prediction location mode, duplicates, regression)."""

    def __init__(self, params):
        self.params = UserParams.from_dict(params)
        d = C.c_void_p()
        check_io(lib().mra_data_load(C.byref(self.params), C.byref(d)))
        self.d = d
        v = DataView()
        check_io(lib().mra_data_get(self.d, C.byref(v)))
        self.n_locations, self.n_obs_raw, self.n_pred_raw = v.n_locations, v.n_obs_raw, v.n_pred_raw
        self.obs_lon, self.obs_lat = _arr(v.obs_lon, v.n_obs, np.float64), _arr(v.obs_lat, v.n_obs, np.float64)
        self.obs_val = _arr(v.obs_val, v.n_obs, np.float64)
        self.obs_region = _arr(v.obs_region, v.n_obs, np.uint64) if v.obs_region else None
        self.pred_lon, self.pred_lat = _arr(v.pred_lon, v.n_pred, np.float64), _arr(v.pred_lat, v.n_pred, np.float64)
        self.pred_region = _arr(v.pred_region, v.n_pred, np.uint64) if v.pred_region else None
        self.bounds = np.array(v.bounds[:])
        self.coefficients = np.array(v.coefficients[:]) if v.regression else None
        self.predicting = bool(v.predicting)

    def partition(self, params=None):
        """mra_partition_build: the tree dict (as partition_program returns)."""
        p = UserParams.from_dict(params) if params is not None else self.params
        P = C.c_void_p()
        check_io(lib().mra_partition_build(C.byref(p), self.d, C.byref(P)))
        try:
            t = TreeDesc()
            check_io(lib().mra_partition_tree(P, C.byref(t)))
            nC, nF, no, nq = t.n_coarse, t.n_finest, t.n_obs, t.n_pred
            oi, pi = u64p(), u64p()
            check_io(lib().mra_partition_indices(P, C.byref(oi), C.byref(pi)))
            counts = np.zeros(t.num_levels, dtype=np.uint64)
            check_io(lib().mra_partition_level_counts(P, _u(counts)))
            tree = dict(M=t.num_levels, J=t.num_partitions, r=t.num_knots, various_j=t.various_j, n_coarse=nC,
                        n_finest=nF, knot_lon=_arr(t.knot_lon, nC * t.num_knots, np.float64),
                        knot_lat=_arr(t.knot_lat, nC * t.num_knots, np.float64),
                        finest_parent=_arr(t.finest_parent, nF, np.uint64),
                        obs_offsets=_arr(t.obs_offsets, nF + 1, np.uint64), obs_lon=_arr(t.obs_lon, no, np.float64),
                        obs_lat=_arr(t.obs_lat, no, np.float64), obs_val=_arr(t.obs_val, no, np.float64),
                        obs_index=_arr(oi, no, np.uint64), level_counts=counts)
            if t.pred_offsets:
                tree.update(pred_offsets=_arr(t.pred_offsets, nF + 1, np.uint64),
                            pred_lon=_arr(t.pred_lon, nq, np.float64), pred_lat=_arr(t.pred_lat, nq, np.float64),
                            pred_index=_arr(pi, nq, np.uint64))
            return tree
        finally:
            lib().mra_partition_destroy(P)

    def dump_structure(self, path, params=None):
        """mra_partition_dump_structure of the partition of this data set."""
        p = UserParams.from_dict(params) if params is not None else self.params
        P = C.c_void_p()
        check_io(lib().mra_partition_build(C.byref(p), self.d, C.byref(P)))
        try:
            check_io(lib().mra_partition_dump_structure(P, str(path).encode()))
        finally:
            lib().mra_partition_destroy(P)

    def close(self):
        if getattr(self, "d", None):
            lib().mra_data_destroy(self.d)
            self.d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_levels(n, J, r):
    return lib().mra_default_levels(n, J, r)


def load_cov_model(params):
    """mra_cov_model_load: the covariance model of a parameter dict as the
    cov dict of Handle.set_covariance (arrays copied)."""
    p = UserParams.from_dict(params)
    m = C.c_void_p()
    check_io(lib().mra_cov_model_load(C.byref(p), C.byref(m)))
    try:
        c = CovDesc()
        check_io(lib().mra_cov_model_desc(m, C.byref(c)))

        def fld(f):
            d = dict(kind=f.kind, intercept=f.intercept, slope_x=f.slope_x, slope_y=f.slope_y)
            if f.kind == FIELD_WENDLAND:
                n = f.n_basis
                d.update(lon=_arr(f.basis_lon, n, np.float64), lat=_arr(f.basis_lat, n, np.float64),
                         w=_arr(f.basis_weight, n, np.float64))
            return d
        return dict(nonstationary=c.nonstationary, metric=c.metric, wendland_len=c.wendland_len,
                    sigmasq=fld(c.sigmasq), beta=fld(c.beta), tausq=fld(c.tausq))
    finally:
        lib().mra_cov_model_destroy(m)


def write_prediction_result(path, lon, lat, mean, var, coef=None):
    keep = Keep()
    n = len(lon)
    check_io(lib().mra_write_prediction_result(str(path).encode(), n, keep.d(lon), keep.d(lat), keep.d(mean),
                                               keep.d(var), keep.d(coef) if coef is not None else None))


def fields(cov, lon=None, lat=None, x=None, y=None, z=None):
    """mra_fields: (sigmasq, beta, tausq) at the points, unscaled."""
    keep = Keep()
    c = cov_desc(cov, keep)
    n = len(lon) if lon is not None else len(x)
    s, b, t = np.zeros(n), np.zeros(n), np.zeros(n)
    opt = lambda a: keep.d(a) if a is not None else None
    check(lib().mra_fields(C.byref(c), opt(lon), opt(lat), opt(x), opt(y), opt(z), n, _d(s), _d(b), _d(t)))
    return s, b, t


def selftest(K, Mv, Nv, A, B, C0, S, Bt):
    """Device building-block self test (see mra_selftest in include/mra.h)."""
    keep = Keep()
    Co, Lo, Xo = np.zeros(64 * 64), np.zeros(64 * 64), np.zeros(64 * 64)
    bad = C.c_int(0)
    check(lib().mra_selftest(K, Mv, Nv, keep.d(A), keep.d(B), keep.d(C0), keep.d(S), keep.d(Bt), _d(Co), _d(Lo),
                             _d(Xo), C.byref(bad)))
    return Co.reshape(64, 64), Lo.reshape(64, 64), Xo.reshape(64, 64), bad.value


OBJECTIVE = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)
MRA_W_MAX_EVALS = 1
TREND_NPARAM = 6


def bobyqa(fun, x0, lo, hi, rho_begin, rho_end, max_evals, npt=0, h0=None):
    """mra_bobyqa on a Python objective fun(x) -> float.  Returns
    (x_best, f_best, n_evals, status) with status MRA_OK or MRA_W_MAX_EVALS."""
    n = len(x0)
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    cb = OBJECTIVE(lambda p, ctx: float(fun(np.ctypeslib.as_array(p, shape=(n,)).copy())))
    fb = C.c_double(0.0)
    ne = C.c_int(0)
    H0 = np.ascontiguousarray(h0, dtype=np.float64).reshape(-1) if h0 is not None else None
    rc = lib().mra_bobyqa_h0(n, npt, cb, None, _d(x), _d(lo), _d(hi), rho_begin, rho_end, max_evals,
                             _d(H0) if H0 is not None else None, C.byref(fb), C.byref(ne))
    if rc < 0:
        check(rc)
    return x, fb.value, ne.value, rc


class Batch:
    """Independent handles evaluated as one launch sequence (mra_batch_*)."""

    def __init__(self, handles):
        self.handles = list(handles)  # keeps the handles alive
        arr = (C.c_void_p * len(self.handles))(*[h.h for h in self.handles])
        b = C.c_void_p()
        check(lib().mra_batch_create(arr, len(self.handles), C.byref(b)))
        self.b = b

    def _check(self, rc):
        if rc != MRA_OK:
            msg = lib().mra_batch_last_error(self.b)
            raise MraError(rc, msg.decode() if msg else "")

    def set(self, streams, graph):
        self._check(lib().mra_batch_set(self.b, streams, int(graph)))

    def get(self):
        s, g = C.c_int(0), C.c_int(0)
        self._check(lib().mra_batch_get(self.b, C.byref(s), C.byref(g)))
        return s.value, g.value

    def likelihood(self):
        n = len(self.handles)
        D = np.zeros(n)
        st = (C.c_int * n)()
        self._check(lib().mra_batch_likelihood(self.b, _d(D), st))
        return D, np.array(st[:], dtype=np.int32)

    def autotune(self):
        s, g, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        self._check(lib().mra_batch_autotune(self.b, C.byref(s), C.byref(g), C.byref(ms)))
        return s.value, g.value, ms.value

    def close(self):
        if getattr(self, "b", None):
            lib().mra_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The warm-start network (mra_mlp_*): 64 -> 256 -> 256 -> 6, GELU."""

    NIN, NOUT = 64, 6

    def __init__(self, seed=0):
        h = C.c_void_p()
        rc = lib().mra_mlp_create(seed, C.byref(h))
        if rc:
            raise MraError(rc, lib().mra_mlp_last_error(None).decode())
        self.h = h
        self.nparam = lib().mra_mlp_param_count()

    def _check(self, rc):
        if rc:
            raise MraError(rc, lib().mra_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().mra_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._check(lib().mra_mlp_get_params(self.h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(lib().mra_mlp_set_params(self.h, _f(p)))

    MAXB = 1024

    def forward(self, X, bf16=False):
        """Outputs for every row of X (in chunks of at most 1024 rows)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.NOUT), dtype=np.float32)
        for b0 in range(0, X.shape[0], self.MAXB):
            xb = np.ascontiguousarray(X[b0:b0 + self.MAXB])
            yb = np.zeros((xb.shape[0], self.NOUT), dtype=np.float32)
            self._check(lib().mra_mlp_forward(self.h, _f(xb), xb.shape[0], _f(yb), int(bf16)))
            Y[b0:b0 + xb.shape[0]] = yb
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float(0)
        self._check(lib().mra_mlp_loss_grad(self.h, _f(X), _f(T), X.shape[0], C.byref(loss), _f(g)))
        return loss.value, g

    def train_step(self, X, T, lr):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float(0)
        self._check(lib().mra_mlp_train_step(self.h, _f(X), _f(T), X.shape[0], lr, C.byref(loss)))
        return loss.value


def selftest_blk(S):
    """Blocked Cholesky / inverse self test (see mra_selftest_blk in include/mra.h)."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    n = S.shape[0]
    Lo, Xo = np.zeros((n, n)), np.zeros((n, n))
    bad = C.c_int(0)
    check(lib().mra_selftest_blk(n, _d(S), _d(Lo), _d(Xo), C.byref(bad)))
    return Lo, Xo, bad.value


def selftest_f32(K, Mv, Nv, A, B, C0):
    """FP32 GEMM self test (see mra_selftest_f32 in include/mra.h)."""
    fp = C.POINTER(C.c_float)
    a = np.ascontiguousarray(A, dtype=np.float32)
    b = np.ascontiguousarray(B, dtype=np.float32)
    c0 = np.ascontiguousarray(C0, dtype=np.float32)
    out = np.zeros(64 * 64, dtype=np.float32)
    check(lib().mra_selftest_f32(K, Mv, Nv, a.ctypes.data_as(fp), b.ctypes.data_as(fp), c0.ctypes.data_as(fp),
                                 out.ctypes.data_as(fp)))
    return out.reshape(64, 64)


def check(rc, h=None):
    if rc != MRA_OK:
        msg = lib().mra_last_error(h)
        raise MraError(rc, msg.decode() if msg else "")


def _d(a):
    return a.ctypes.data_as(dp)


def _u(a):
    return a.ctypes.data_as(u64p)


class Keep:
    def __init__(self):
        self.refs = []

    def d(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        self.refs.append(a)
        return _d(a)

    def u(self, x):
        a = np.ascontiguousarray(x, dtype=np.uint64)
        self.refs.append(a)
        return _u(a)


def tree_desc(tree, keep, with_pred=True):
    t = TreeDesc()
    t.num_levels, t.num_partitions, t.num_knots = tree["M"], tree["J"], tree["r"]
    t.various_j = int(tree.get("various_j", 0))
    t.n_coarse, t.n_finest = tree["n_coarse"], tree["n_finest"]
    t.n_obs = len(tree["obs_lon"])
    t.knot_lon, t.knot_lat = keep.d(tree["knot_lon"]), keep.d(tree["knot_lat"])
    t.finest_parent = keep.u(tree["finest_parent"])
    t.obs_offsets = keep.u(tree["obs_offsets"])
    t.obs_lon, t.obs_lat, t.obs_val = keep.d(tree["obs_lon"]), keep.d(tree["obs_lat"]), keep.d(tree["obs_val"])
    if with_pred and tree.get("pred_offsets") is not None and len(tree["pred_lon"]):
        t.n_pred = len(tree["pred_lon"])
        t.pred_offsets = keep.u(tree["pred_offsets"])
        t.pred_lon, t.pred_lat = keep.d(tree["pred_lon"]), keep.d(tree["pred_lat"])
    return t


def _field(f, keep):
    o = FieldDesc()
    o.kind = f["kind"]
    o.intercept = f.get("intercept", 0.0)
    if f["kind"] == FIELD_WENDLAND:
        o.n_basis = len(f["lon"])
        o.basis_lon, o.basis_lat, o.basis_weight = keep.d(f["lon"]), keep.d(f["lat"]), keep.d(f["w"])
    o.slope_x = f.get("slope_x", 0.0)
    o.slope_y = f.get("slope_y", 0.0)
    return o


def cov_desc(cov, keep):
    c = CovDesc()
    c.nonstationary = int(cov["nonstationary"])
    c.metric = cov["metric"]
    c.wendland_len = cov.get("wendland_len", 1.0)
    c.sigmasq = _field(cov["sigmasq"], keep)
    c.beta = _field(cov["beta"], keep)
    c.tausq = _field(cov["tausq"], keep)
    return c


class Handle:
    """Owns one mra_handle (device-resident MRA structure)."""

    def __init__(self, tree, with_pred=True):
        keep = Keep()
        t = tree_desc(tree, keep, with_pred)
        h = C.c_void_p()
        check(lib().mra_create(C.byref(t), C.byref(h)))
        self.h = h
        self.n_obs = int(t.n_obs)
        self.n_pred = int(t.n_pred)
        self.M = tree["M"]

    def fit_linear_trend(self, theta0, lo, hi, rho_begin=0.1, rho_end=1e-6, max_evals=400, metric=METRIC_PLANAR,
                         npt=0, h0=None):
        """Maximum-likelihood fit of the six linear-trend parameters
        (mra_fit_linear_trend_npt; h0 = starting curvature or None).
        Returns (theta, D, n_evals, status)."""
        th = np.ascontiguousarray(theta0, dtype=np.float64).copy()
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        D = C.c_double(0.0)
        ne = C.c_int(0)
        H0 = np.ascontiguousarray(h0, dtype=np.float64).reshape(-1) if h0 is not None else None
        rc = lib().mra_fit_linear_trend_npt(self.h, metric, _d(th), _d(lo), _d(hi), rho_begin, rho_end, max_evals,
                                            npt, _d(H0) if H0 is not None else None, C.byref(D), C.byref(ne))
        if rc < 0:
            check(rc, self.h)
        return th, D.value, ne.value, rc

    def set_tuning(self, x32max, lrmax, fb16):
        check(lib().mra_set_tuning(self.h, x32max, lrmax, fb16), self.h)

    def get_tuning(self):
        v = [C.c_int(0) for _ in range(3)]
        check(lib().mra_get_tuning(self.h, *[C.byref(x) for x in v]), self.h)
        return tuple(x.value for x in v)

    def autotune(self):
        """mra_autotune: returns ((x32max, lrmax, fb16), best_ms)."""
        v = [C.c_int(0) for _ in range(3)]
        ms = C.c_double(0)
        check(lib().mra_autotune(self.h, *[C.byref(x) for x in v], C.byref(ms)), self.h)
        return tuple(x.value for x in v), ms.value

    def variogram_features(self):
        """64 log-semivariogram features of the current observation values
        (mra_variogram_features; planar metric)."""
        f = np.zeros(64, dtype=np.float32)
        check(lib().mra_variogram_features(self.h, _f(f)), self.h)
        return f

    def close(self):
        if getattr(self, "h", None):
            lib().mra_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_covariance(self, cov):
        keep = Keep()
        c = cov_desc(cov, keep)
        check(lib().mra_set_covariance(self.h, C.byref(c)), self.h)

    def set_params(self, sill=0.0, range=1.0, nugget=0.0, sill_scale=1.0, range_scale=1.0, nugget_scale=1.0):
        check(lib().mra_set_params(self.h, sill, range, nugget, sill_scale, range_scale, nugget_scale), self.h)

    def set_precision(self, p):
        check(lib().mra_set_precision(self.h, p), self.h)

    def set_option(self, option, value):
        check(lib().mra_set_option(self.h, option, int(value)), self.h)

    def get_option(self, option):
        v = C.c_int()
        check(lib().mra_get_option(self.h, option, C.byref(v)), self.h)
        return v.value

    def use_graph(self, on=True):
        check(lib().mra_use_graph(self.h, int(on)), self.h)

    def likelihood(self):
        D = C.c_double()
        check(lib().mra_likelihood(self.h, C.byref(D)), self.h)
        return D.value

    def loglikelihood(self, n_observations=0):
        """mra_loglikelihood: (D, -0.5 (D + n log(2*3.14159265359)))."""
        D, ell = C.c_double(), C.c_double()
        check(lib().mra_loglikelihood(self.h, n_observations, C.byref(D), C.byref(ell)), self.h)
        return D.value, ell.value

    def likelihood_levels(self):
        out = np.zeros(self.M)
        check(lib().mra_likelihood_levels(self.h, _d(out)), self.h)
        return out

    def predict(self):
        D = C.c_double()
        mean, var = np.zeros(self.n_pred), np.zeros(self.n_pred)
        check(lib().mra_predict(self.h, C.byref(D), _d(mean), _d(var)), self.h)
        return D.value, mean, var

    def sample_prior(self, seed, replace=True):
        y = np.zeros(self.n_obs)
        check(lib().mra_sample_prior(self.h, seed, _d(y), int(replace)), self.h)
        return y

    def set_observations(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        check(lib().mra_set_observations(self.h, _d(y)), self.h)

    def timings(self):
        """ms per launch group of the last evaluation (index = MRA_T_* ids)."""
        out = np.zeros(T_COUNT)
        check(lib().mra_last_timings(self.h, _d(out), T_COUNT), self.h)
        return out


def cross_covariance(cov, prm, lon1, lat1, lon2, lat2):
    keep = Keep()
    c = cov_desc(cov, keep)
    n1, n2 = len(lon1), len(lon2)
    out = np.zeros(n1 * n2)
    check(lib().mra_cross_covariance(C.byref(c), prm.get("sill", 0.0), prm.get("range", 1.0), prm.get("nugget", 0.0),
                                     prm.get("sill_scale", 1.0), prm.get("range_scale", 1.0),
                                     prm.get("nugget_scale", 1.0), keep.d(lon1), keep.d(lat1), n1, keep.d(lon2),
                                     keep.d(lat2), n2, _d(out)))
    return out.reshape(n1, n2)


def variance_covariance(cov, prm, lon, lat):
    keep = Keep()
    c = cov_desc(cov, keep)
    n = len(lon)
    out = np.zeros(n * n)
    check(lib().mra_variance_covariance(C.byref(c), prm.get("sill", 0.0), prm.get("range", 1.0),
                                        prm.get("nugget", 0.0), prm.get("sill_scale", 1.0),
                                        prm.get("range_scale", 1.0), prm.get("nugget_scale", 1.0), keep.d(lon),
                                        keep.d(lat), n, _d(out)))
    return out.reshape(n, n)


class DeviceBuffer:
    """A device allocation made through the library (mra_device_alloc), so
    that callers need no HIP runtime of their own."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().mra_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().mra_device_copy(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))
        return b

    def to_array(self, dtype, count, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        assert offset_bytes + out.nbytes <= self.nbytes
        check(lib().mra_device_copy(out.ctypes.data_as(C.c_void_p), self.ptr + offset_bytes, out.nbytes, 0))
        return out

    def upload(self, a, offset_bytes=0):
        a = np.ascontiguousarray(a)
        assert offset_bytes + a.nbytes <= self.nbytes
        check(lib().mra_device_copy(self.ptr + offset_bytes, a.ctypes.data_as(C.c_void_p), a.nbytes, 1))

    def free(self):
        if self.ptr:
            check(lib().mra_device_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cov_matrix_device(precision, metric, xyz1, sb1, n1, xyz2, sb2, n2, out, stream=None):
    """Device covariance matrix out[i*n2 + j] (mra_cov_tile_device).  The
    arguments are device addresses (ints): xyz = [x | y | z] and sb = [s | b]
    (scaled sqrt-sill and range) as FP64 SoA, out FP64 or, for
    PRECISION_MIXED, FP32; the launch is asynchronous on `stream`."""
    check(lib().mra_cov_tile_device(precision, metric, xyz1, sb1, n1, xyz2, sb2, n2, out, stream))


def cov_matrix_timed(precision, metric, xyz1, sb1, n1, xyz2, sb2, n2, out, launches):
    """Average milliseconds per launch of mra_cov_tile_device (HIP events on
    the library's own stream, after one untimed launch)."""
    ms = C.c_double()
    check(lib().mra_cov_tile_timed(precision, metric, xyz1, sb1, n1, xyz2, sb2, n2, out, launches, C.byref(ms)))
    return ms.value


def partition_program(lon, lat, val, plon, plat, M, J, r, offset=np.e / 100, bounds=None, gpu=False):
    """Program-placed partition: the library's host builder, or with gpu=True
    its GPU builder (mra_partition_program_gpu; identical results)."""
    if gpu:
        return partition_program_gpu(lon, lat, val, plon, plat, M, J, r, offset, bounds)
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    plon = np.ascontiguousarray(plon if plon is not None else np.zeros(0), dtype=np.float64)
    plat = np.ascontiguousarray(plat if plat is not None else np.zeros(0), dtype=np.float64)
    if bounds is None:
        mnl, mxl, mna, mxa = lon.min(), lon.max(), lat.min(), lat.max()
        bounds = np.array([mnl, mxl + 1e-6 * (mxl - mnl), mna, mxa + 1e-6 * (mxa - mna)])
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    n, npred = len(lon), len(plon)
    sizes = np.zeros(4, dtype=np.uint64)
    check(lib().mra_partition_program_sizes(_d(lon), _d(lat), n, _d(plon), _d(plat), npred, M, J, r, offset, _d(b),
                                            _u(sizes)))
    nk, nF, no, nq = (int(x) for x in sizes)
    nC = sum(J ** l for l in range(M - 1))
    kl, kt = np.zeros(nk), np.zeros(nk)
    parent = np.zeros(nF, dtype=np.uint64)
    oo = np.zeros(nF + 1, dtype=np.uint64)
    po = np.zeros(nF + 1, dtype=np.uint64)
    ol, ot, ov = np.zeros(max(no, 1)), np.zeros(max(no, 1)), np.zeros(max(no, 1))
    oi = np.zeros(max(no, 1), dtype=np.uint64)
    pl, pt = np.zeros(max(nq, 1)), np.zeros(max(nq, 1))
    pi = np.zeros(max(nq, 1), dtype=np.uint64)
    check(lib().mra_partition_program(_d(lon), _d(lat), _d(val), n, _d(plon), _d(plat), npred, M, J, r, offset,
                                      _d(b), _d(kl), _d(kt), _u(parent), _u(oo), _d(ol), _d(ot), _d(ov), _u(oi),
                                      _u(po), _d(pl), _d(pt), _u(pi)))
    return dict(M=M, J=J, r=nk // nC, various_j=0, n_coarse=nC, n_finest=nF, knot_lon=kl, knot_lat=kt,
                finest_parent=parent, obs_offsets=oo, obs_lon=ol[:no], obs_lat=ot[:no], obs_val=ov[:no],
                obs_index=oi[:no], pred_offsets=po, pred_lon=pl[:nq], pred_lat=pt[:nq], pred_index=pi[:nq], bounds=b)


def partition_program_gpu(lon, lat, val, plon, plat, M, J, r, offset=np.e / 100, bounds=None):
    """mra_partition_program_gpu: the program partition built on the GPU."""
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    plon = np.ascontiguousarray(plon if plon is not None else np.zeros(0), dtype=np.float64)
    plat = np.ascontiguousarray(plat if plat is not None else np.zeros(0), dtype=np.float64)
    if bounds is None:
        mnl, mxl, mna, mxa = lon.min(), lon.max(), lat.min(), lat.max()
        bounds = np.array([mnl, mxl + 1e-6 * (mxl - mnl), mna, mxa + 1e-6 * (mxa - mna)])
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    n, npred = len(lon), len(plon)
    nC = sum(J ** l for l in range(M - 1))
    nlon = int(np.ceil(np.sqrt(r)))
    nk = nC * nlon * (r // nlon)
    nF = J ** (M - 1)
    kl, kt = np.zeros(nk), np.zeros(nk)
    parent = np.zeros(nF, dtype=np.uint64)
    oo, po = np.zeros(nF + 1, dtype=np.uint64), np.zeros(nF + 1, dtype=np.uint64)
    ol, ot, ov = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
    oi = np.zeros(max(n, 1), dtype=np.uint64)
    pl, pt = np.zeros(max(npred, 1)), np.zeros(max(npred, 1))
    pi = np.zeros(max(npred, 1), dtype=np.uint64)
    sizes = np.zeros(4, dtype=np.uint64)
    rc = lib().mra_partition_program_gpu(_d(lon), _d(lat), _d(val), n, _d(plon), _d(plat), npred, M, J, r, offset,
                                         _d(b), _d(kl), _d(kt), _u(parent), _u(oo), _d(ol), _d(ot), _d(ov), _u(oi),
                                         _u(po), _d(pl), _d(pt), _u(pi), _u(sizes))
    if rc != MRA_OK:
        raise MraError(rc, lib().mra_partition_gpu_last_error().decode())
    no, nq = int(sizes[2]), int(sizes[3])
    return dict(M=M, J=J, r=nk // nC, various_j=0, n_coarse=nC, n_finest=nF, knot_lon=kl, knot_lat=kt,
                finest_parent=parent, obs_offsets=oo, obs_lon=ol[:no], obs_lat=ot[:no], obs_val=ov[:no],
                obs_index=oi[:no], pred_offsets=po, pred_lon=pl[:nq], pred_lat=pt[:nq], pred_index=pi[:nq], bounds=b)
