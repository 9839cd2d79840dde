"""ctypes binding of the CPU FP64 oracle (oracle/mra_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker the HIP product path is compared
against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmra_oracle.so")

u64p = C.POINTER(C.c_uint64)
dp = C.POINTER(C.c_double)


class OrcTree(C.Structure):
    _fields_ = [("num_levels", C.c_int32), ("num_partitions", C.c_int32), ("num_knots", C.c_int32),
                ("various_j", C.c_int32), ("n_coarse", C.c_uint64), ("n_finest", C.c_uint64),
                ("n_obs", C.c_uint64), ("n_pred", C.c_uint64), ("knot_lon", dp), ("knot_lat", dp),
                ("finest_parent", u64p), ("obs_offsets", u64p), ("obs_lon", dp), ("obs_lat", dp),
                ("obs_val", dp), ("pred_offsets", u64p), ("pred_lon", dp), ("pred_lat", dp)]


class OrcField(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("n_basis", C.c_uint64), ("intercept", C.c_double),
                ("basis_lon", dp), ("basis_lat", dp), ("basis_weight", dp), ("slope_x", C.c_double),
                ("slope_y", C.c_double)]


class OrcCov(C.Structure):
    _fields_ = [("nonstationary", C.c_int32), ("metric", C.c_int32), ("wendland_len", C.c_double),
                ("sigmasq", OrcField), ("beta", OrcField), ("tausq", OrcField)]


class OrcParams(C.Structure):
    _fields_ = [("sill", C.c_double), ("range", C.c_double), ("nugget", C.c_double),
                ("sill_scale", C.c_double), ("range_scale", C.c_double), ("nugget_scale", C.c_double)]


def build(force=False):
    """Compile the oracle with g++ into oracle/_build/."""
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    src = os.path.join(HERE, "mra_oracle.cpp")
    if not force and os.path.exists(LIB_PATH) and os.path.getmtime(LIB_PATH) > os.path.getmtime(src):
        return LIB_PATH
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", src, "-o", LIB_PATH + ".tmp", "-ldl"]
    subprocess.check_call(cmd)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_evaluate.restype = C.c_int
        _lib.orc_set_blas.restype = C.c_int
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _u(a):
    return a.ctypes.data_as(u64p)


class _Keep:
    """Holds numpy arrays alive while ctypes structs point at them."""

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


def make_tree(tree, keep, with_pred=True):
    t = OrcTree()
    t.num_levels, t.num_partitions, t.num_knots = tree["M"], tree["J"], tree["r"]
    t.various_j = int(tree.get("various_j", 0))
    t.n_coarse, t.n_finest = tree["n_coarse"], tree["n_finest"]
    t.n_obs = len(tree["obs_lon"])
    t.knot_lon, t.knot_lat = keep.d(tree["knot_lon"]), keep.d(tree["knot_lat"])
    t.finest_parent = keep.u(tree["finest_parent"])
    t.obs_offsets = keep.u(tree["obs_offsets"])
    t.obs_lon, t.obs_lat, t.obs_val = keep.d(tree["obs_lon"]), keep.d(tree["obs_lat"]), keep.d(tree["obs_val"])
    if with_pred and tree.get("pred_offsets") is not None:
        t.n_pred = len(tree["pred_lon"])
        t.pred_offsets = keep.u(tree["pred_offsets"])
        t.pred_lon, t.pred_lat = keep.d(tree["pred_lon"]), keep.d(tree["pred_lat"])
    else:
        t.n_pred = 0
    return t


def _field(f, keep):
    o = OrcField()
    o.kind = f["kind"]
    o.intercept = f.get("intercept", 0.0)
    if f["kind"] == 1:
        o.n_basis = len(f["lon"])
        o.basis_lon, o.basis_lat, o.basis_weight = keep.d(f["lon"]), keep.d(f["lat"]), keep.d(f["w"])
    o.slope_x = f.get("slope_x", 0.0)
    o.slope_y = f.get("slope_y", 0.0)
    return o


def make_cov(cov, keep):
    c = OrcCov()
    c.nonstationary = int(cov["nonstationary"])
    c.metric = cov["metric"]
    c.wendland_len = cov.get("wendland_len", 1.0)
    c.sigmasq = _field(cov["sigmasq"], keep)
    c.beta = _field(cov["beta"], keep)
    c.tausq = _field(cov["tausq"], keep)
    return c


def make_params(p):
    return OrcParams(p.get("sill", 0.0), p.get("range", 1.0), p.get("nugget", 0.0), p.get("sill_scale", 1.0),
                     p.get("range_scale", 1.0), p.get("nugget_scale", 1.0))


def evaluate(tree, cov, prm, predict=True, nthreads=0):
    """Returns (D, mean, var) — mean/var in finest-region order or None."""
    keep = _Keep()
    t = make_tree(tree, keep, with_pred=predict)
    c = make_cov(cov, keep)
    p = make_params(prm)
    D = C.c_double(0)
    mean = var = None
    if predict and t.n_pred:
        mean = np.zeros(t.n_pred)
        var = np.zeros(t.n_pred)
        rc = lib().orc_evaluate(C.byref(t), C.byref(c), C.byref(p), C.byref(D), _d(mean), _d(var), nthreads)
    else:
        rc = lib().orc_evaluate(C.byref(t), C.byref(c), C.byref(p), C.byref(D), None, None, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle evaluate failed rc={rc}")
    return D.value, mean, var


def evaluate_timed(tree, cov, prm, nthreads=0):
    """Returns (D, phase seconds): coarse prior per level, finest level,
    posterior per level (2 M - 1 values; measurement only)."""
    keep = _Keep()
    t = make_tree(tree, keep, with_pred=False)
    c = make_cov(cov, keep)
    p = make_params(prm)
    D = C.c_double(0)
    ph = np.zeros(2 * tree["M"] - 1)
    rc = lib().orc_evaluate_timed(C.byref(t), C.byref(c), C.byref(p), C.byref(D), _d(ph), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle evaluate_timed failed rc={rc}")
    return D.value, ph


def evaluate_levels(tree, cov, prm, nthreads=0):
    keep = _Keep()
    t = make_tree(tree, keep, with_pred=False)
    c = make_cov(cov, keep)
    p = make_params(prm)
    out = np.zeros(tree["M"])
    rc = lib().orc_evaluate_levels(C.byref(t), C.byref(c), C.byref(p), _d(out), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle evaluate_levels failed rc={rc}")
    return out


def cross_covariance(cov, prm, lon1, lat1, lon2, lat2):
    keep = _Keep()
    c = make_cov(cov, keep)
    p = make_params(prm)
    n1, n2 = len(lon1), len(lon2)
    out = np.zeros(n1 * n2)
    lib().orc_cross_covariance(C.byref(c), C.byref(p), keep.d(lon1), keep.d(lat1), C.c_uint64(n1), keep.d(lon2),
                               keep.d(lat2), C.c_uint64(n2), _d(out))
    return out.reshape(n1, n2)


def variance_covariance(cov, prm, lon, lat):
    keep = _Keep()
    c = make_cov(cov, keep)
    p = make_params(prm)
    n = len(lon)
    out = np.zeros(n * n)
    lib().orc_variance_covariance(C.byref(c), C.byref(p), keep.d(lon), keep.d(lat), C.c_uint64(n), _d(out))
    return out.reshape(n, n)


def fields(cov, lon, lat):
    keep = _Keep()
    c = make_cov(cov, keep)
    n = len(lon)
    s, b, t = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_fields(C.byref(c), keep.d(lon), keep.d(lat), C.c_uint64(n), _d(s), _d(b), _d(t))
    return s, b, t


def domain_bounds(lon, lat):
    """Data::load_data bounds with the 1e-6 pad (src/class_data.cpp:149-163)."""
    lon = np.asarray(lon)
    lat = np.asarray(lat)
    mnl, mxl, mna, mxa = lon.min(), lon.max(), lat.min(), lat.max()
    return np.array([mnl, mxl + 1e-6 * (mxl - mnl), mna, mxa + 1e-6 * (mxa - mna)])


def partition_program(lon, lat, val, plon, plat, M, J, r, offset=np.e / 100, bounds=None):
    """Restated build_partition_by_program (src/class_partition.cpp:459-654).

    The reference scans every observation for every finest region (O(n*F));
    only small inputs should be passed here."""
    keep = _Keep()
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    plon = np.ascontiguousarray(plon if plon is not None else np.zeros(0), dtype=np.float64)
    plat = np.ascontiguousarray(plat if plat is not None else np.zeros(0), dtype=np.float64)
    if bounds is None:
        bounds = domain_bounds(lon, lat)
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    n, npred = len(lon), len(plon)
    nlon = int(np.ceil(np.sqrt(r)))
    rk = nlon * (r // nlon)
    nC = sum(J ** l for l in range(M - 1))
    nF = J ** (M - 1)
    kl, kt = np.zeros(nC * rk), np.zeros(nC * rk)
    oo = np.zeros(nF + 1, dtype=np.uint64)
    po = np.zeros(nF + 1, dtype=np.uint64)
    ol, ot, ov = np.zeros(n), np.zeros(n), np.zeros(n)
    oi = np.zeros(n, dtype=np.uint64)
    pl, pt = np.zeros(max(npred, 1)), np.zeros(max(npred, 1))
    pi = np.zeros(max(npred, 1), dtype=np.uint64)
    lib().orc_partition_program(_d(lon), _d(lat), _d(val), C.c_uint64(n), _d(plon), _d(plat), C.c_uint64(npred),
                                M, J, r, C.c_double(offset), _d(b), _d(kl), _d(kt), _u(oo), _d(ol), _d(ot), _d(ov),
                                _u(oi), _u(po), _d(pl), _d(pt), _u(pi))
    no, nq = int(oo[-1]), int(po[-1])
    parent = np.array([(nC + f - 1) // J for f in range(nF)], dtype=np.uint64)
    del keep
    return dict(M=M, J=J, r=rk, various_j=0, n_coarse=nC, n_finest=nF, knot_lon=kl, knot_lat=kt,
                finest_parent=parent, obs_offsets=oo, obs_lon=ol[:no], obs_lat=ot[:no], obs_val=ov[:no],
                obs_index=oi[:no], pred_offsets=po, pred_lon=pl[:nq], pred_lat=pt[:nq], pred_index=pi[:nq],
                bounds=b)


def partition_user(knots, parent_file, obs, pred, M, J, r):
    """Restated build_partition_by_user_files + read_parent_region_index
    (src/class_partition.cpp:93-454, 659-770).

    knots: dict(lon, lat, region) as read from the knots file (grouped r per
    coarse region); parent_file: dict(start, end, parent); obs: dict(lon,
    lat, val, region); pred: dict(lon, lat, region) or None."""
    nC = sum(J ** l for l in range(M - 1))
    start, end = int(parent_file["start"]), int(parent_file["end"])
    if start != nC:
        raise ValueError("finest region ids must start right after the coarse regions")
    nF = end - start + 1
    kl, kt = np.zeros(nC * r), np.zeros(nC * r)
    kr = np.asarray(knots["region"], dtype=np.int64)
    i = 0
    while i < len(kr):
        if kr[i] < nC:
            R = kr[i]
            kl[R * r:(R + 1) * r] = knots["lon"][i:i + r]
            kt[R * r:(R + 1) * r] = knots["lat"][i:i + r]
            i += r
        else:
            break

    def group(region, n):
        reg = np.asarray(region, dtype=np.int64)
        order = []
        offs = np.zeros(nF + 1, dtype=np.uint64)
        for f in range(nF):
            idx = np.nonzero(reg == start + f)[0]
            order.append(idx)
            offs[f + 1] = offs[f] + len(idx)
        return offs, (np.concatenate(order) if order else np.zeros(0, dtype=np.int64))

    oo, oidx = group(obs["region"], len(obs["lon"]))
    tree = dict(M=M, J=J, r=r, various_j=1, n_coarse=nC, n_finest=nF, knot_lon=kl, knot_lat=kt,
                finest_parent=np.asarray(parent_file["parent"], dtype=np.uint64), obs_offsets=oo,
                obs_lon=np.asarray(obs["lon"])[oidx], obs_lat=np.asarray(obs["lat"])[oidx],
                obs_val=np.asarray(obs["val"])[oidx], obs_index=oidx.astype(np.uint64))
    if pred is not None:
        po, pidx = group(pred["region"], len(pred["lon"]))
        tree.update(pred_offsets=po, pred_lon=np.asarray(pred["lon"])[pidx], pred_lat=np.asarray(pred["lat"])[pidx],
                    pred_index=pidx.astype(np.uint64))
    return tree
