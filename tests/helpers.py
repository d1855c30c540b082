"""Shared test plumbing: the CPU oracle (oracle/liboracle.so), the engine's
ctypes module, the reference's golden fixtures and synthetic inputs."""
import ctypes
import importlib.util
import os
import struct

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG = os.path.join(ROOT, "semper-dic-engine_amd")
VP = ctypes.c_void_p


def P(a):
    return a.ctypes.data_as(VP) if a is not None else None


def engine_module():
    spec = importlib.util.spec_from_file_location("semper_amd", os.path.join(PKG, "semper_amd.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_oracle = None


def oracle():
    """The CPU oracle (built by __graft_entry__.build() / `make -C oracle`)."""
    global _oracle
    if _oracle is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built")
        lib = ctypes.CDLL(path)
        lib.orc_inc_create.restype = VP
        lib.orc_inc_destroy.argtypes = [VP]
        lib.orc_inc_init.argtypes = [VP, VP, ctypes.c_int, ctypes.c_int, VP, VP, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float]
        lib.orc_inc_process.argtypes = [VP, VP, VP, ctypes.c_int, VP, ctypes.c_int, VP]
        lib.orc_engine_create.restype = VP
        lib.orc_engine_destroy.argtypes = [VP]
        lib.orc_engine_set_reference.argtypes = [VP, VP, ctypes.c_int, ctypes.c_int, VP]
        lib.orc_engine_run.argtypes = [VP, VP, VP, VP, VP, ctypes.c_int, VP, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float]
        lib.orc_engine_run_planes.argtypes = [VP, VP, VP, VP, VP, ctypes.c_int, VP, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_float, VP, ctypes.c_int, VP]
        lib.orc_solve_nodes.argtypes = [VP, VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_int, VP, VP, VP, VP, VP]
        lib.orc_icgn_fp64.argtypes = [VP, VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      VP, VP, ctypes.c_double, ctypes.c_int, VP, VP, VP]
        lib.orc_ref_speckle.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int, VP, VP]
        lib.orc_gray8.argtypes = [VP, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, VP]
        lib.orc_golden_corpus.argtypes = [ctypes.c_int, ctypes.c_int, VP, VP]
        lib.orc_pool.argtypes = [VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, VP, VP, VP, VP, VP]
        lib.orc_gradients.argtypes = [VP, ctypes.c_int, ctypes.c_int, VP, VP]
        lib.orc_interp.argtypes = [VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, VP, VP, VP]
        lib.orc_outliers_and_vsg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, VP, VP, VP, VP,
                                             VP, VP]
        lib.orc_vsg_window_node_count.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.orc_blob_eval.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_int, VP, VP, VP]
        lib.orc_global_motion.argtypes = [VP, VP, ctypes.c_int, ctypes.c_int, VP, VP]
        lib.orc_hash_u01.restype = ctypes.c_float
        lib.orc_hash_u01.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _oracle = lib
    return _oracle


# ---- the reference's fixtures ------------------------------------------------

CORPUS_DT = np.dtype([("x", "<i4"), ("y", "<i4"), ("status", "<i4"), ("u", "<f4"), ("v", "<f4"), ("ux", "<f4"),
                      ("uy", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("corr", "<f4")])
CORPUS_SCENARIOS = [(3.2, -1.7, 0.0, 0.0, 0.0, 0.0), (2.1, 1.4, 0.015, -0.008, 0.006, 0.012)]


def golden_corpus(kind):
    b = open(os.path.join(GOLDEN, f"golden_corpus.bin.{kind}"), "rb").read()
    n = struct.unpack("<I", b[:4])[0]
    a = np.frombuffer(b[4:], dtype=CORPUS_DT)
    assert len(a) == n
    return a


def full_field_golden():
    b = open(os.path.join(GOLDEN, "full_field_golden.bin"), "rb").read()
    off = 4
    runs = []
    for _ in range(struct.unpack_from("<I", b, 0)[0]):
        code = struct.unpack_from("<i", b, off)[0]
        off += 4
        counts = np.frombuffer(b[off:off + 60], dtype="<f4").copy()
        off += 60
        n = struct.unpack_from("<I", b, off)[0]
        off += 4
        pts = np.frombuffer(b[off:off + 4 * n], dtype="<f4").reshape(-1, 8).copy()
        off += 4 * n
        runs.append((code, counts, pts))
    return runs


FF_SLOTS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 16, 19, 20, 21, 22, 23]
FF_SCENARIOS = [(3.0, 2.0, 0.0, 0.0, 0.0, 0.0), (2.0, -1.5, 0.012, -0.006, 0.005, 0.010)]


# ---- synthetic inputs ----------------------------------------------------------

def ref_speckle(seed, w, h, blobs, aff=None):
    """tests/framework/synthetic.h: float image of the reference's speckle
    field; aff = (u, v, ux, uy, vx, vy, cx, cy) renders the deformed image."""
    out = np.zeros(w * h, np.float32)
    a = None if aff is None else np.array(aff, np.float32)
    oracle().orc_ref_speckle(seed, w, h, blobs, P(a), P(out))
    return out.reshape(h, w)


def gray8(img, sigma=0.0, seed=0):
    h, w = img.shape
    src = np.ascontiguousarray(img, np.float32)
    out = np.zeros(w * h, np.uint8)
    oracle().orc_gray8(P(src), w, h, ctypes.c_float(sigma), seed, P(out))
    return out.reshape(h, w)


def corpus_images(scenario):
    W = H = 512
    ref = ref_speckle(42, W, H, 400)
    s = CORPUS_SCENARIOS[scenario]
    dfm = ref_speckle(42, W, H, 400, list(s) + [W / 2, H / 2])
    return ref, dfm


def corpus_nodes():
    return [(x, y) for y in range(60, 453, 22) for x in range(60, 453, 22)]


def ff_golden_images(scenario):
    W = H = 320
    ref = gray8(ref_speckle(11, W, H, 600))
    s = FF_SCENARIOS[scenario]
    dfm = gray8(ref_speckle(11, W, H, 600, list(s) + [160, 160]))
    return ref, dfm


class OracleEngine:
    """The CPU restatement of semper_run (oracle/src/pipeline.cpp)."""

    def __init__(self):
        self.lib = oracle()
        self.h = self.lib.orc_engine_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_engine_destroy(self.h)

    def set_reference(self, img, mask=None):
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        return self.lib.orc_engine_set_reference(self.h, P(img), w, h, P(m))

    def run(self, img, prm, mask=None, threads=0, tol=1e-3, capacity=None):
        img = np.ascontiguousarray(img, np.uint8)
        pr = np.array(prm, np.int32)
        gw, gh = pr[2] // max(pr[4], 1) + 1, pr[3] // max(pr[4], 1) + 1
        cap = capacity if capacity is not None else gw * gh * 8
        out = np.zeros(max(cap, 1), np.float32)
        met = np.zeros(24, np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        n = self.lib.orc_engine_run(self.h, P(img), P(m), P(pr), P(out), cap, P(met), 24, threads,
                                    ctypes.c_float(tol))
        return n, out[: max(n, 0) * 8].reshape(-1, 8).copy(), met


def _oracle_run_planes(self, img, prm, threads=0, tol=1e-3):
    """run() that also returns the node planes (7, gh, gw): u, v, ux, uy, vx,
    vy, corr (-1 where the node holds no accepted result)."""
    img = np.ascontiguousarray(img, np.uint8)
    pr = np.array(prm, np.int32)
    gw, gh = pr[2] // max(pr[4], 1) + 1, pr[3] // max(pr[4], 1) + 1
    cap = gw * gh * 8
    out = np.zeros(max(cap, 1), np.float32)
    met = np.zeros(24, np.float32)
    planes = np.zeros(7 * gw * gh, np.float32)
    gwh = np.zeros(2, np.int32)
    n = self.lib.orc_engine_run_planes(self.h, P(img), None, P(pr), P(out), cap, P(met), 24, threads,
                                       ctypes.c_float(tol), P(planes), planes.size, P(gwh))
    gw, gh = int(gwh[0]), int(gwh[1])
    return n, out[: max(n, 0) * 8].reshape(-1, 8).copy(), met, planes[: 7 * gw * gh].reshape(7, gh, gw).copy()


OracleEngine.run_planes = _oracle_run_planes


def split_analysis(ref, dfm, prm, planes_g, planes_o, dim, tol, lm=True, agree=1e-5, guess_g=None, path_g=None,
                   tag_g=None, agree_guess=1e-4):
    """Node-by-node comparison of two full-field solves of the same frame
    (planes (7, gh, gw): u, v, ux, uy, vx, vy, corr; corr < 0 unsolved), the
    first side's against the oracle's.

    Both sides iterate IC-GN until ||dp|| < tol, with FP32 arithmetic in
    different orders, so where the two differ by more than `agree` px the
    difference has one of these causes, and this names which:

    "gpu" / "oracle": a stop split. A node whose ||dp|| lands next to tol
    stops one iteration earlier on one side than on the other. IC-GN with a
    fixed Hessian is memoryless -- the next step depends on the current warp
    alone -- so one IC-GN step of the oracle from the earlier side's result
    lands on the later side's result ("gpu": the first side stopped first).

    "guess": the node started from another guess. The anchors' last-bit
    differences (or a stop split at an anchor) move the mesh guesses of the
    Path A nodes around them. With the first side's guess planes (6, gh, gw),
    node paths and mesh tags (sdk_engine_node_report: path 3 = Path A; tags 2
    extrapolated and 4 global model are solved NO_SIMPLEX, others NO_SEARCH),
    the oracle's own solve from that guess lands within `agree_guess` px of
    the first side's result, or one IC-GN step from it. (Two FP32 IC-GN
    trajectories from one guess drift apart by the rounding of their sums
    over their iterations: up to ~6e-5 px over 20 iterations on the
    reference's noiseless affine fixture.)

    "converged": the first side's result is a converged solution in the
    oracle's own arithmetic: the oracle's IC-GN started from it stops at its
    first step (||dp|| < tol) with status 0.

    "-": none of these (a real discrepancy).

    Returns the solved-node masks, the per-node u, v difference, the nodes
    over `agree` px and for each the cause (how) and whether it is a split."""
    vg = planes_g[6] >= 0
    vo = planes_o[6] >= 0
    d = np.maximum(np.abs(planes_g[0] - planes_o[0]), np.abs(planes_g[1] - planes_o[1]))
    d = np.where(vg & vo, d, 0.0)
    iy, ix = np.nonzero(d > agree)
    res = dict(valid_g=vg, valid_o=vo, duv=d, nodes=list(zip(ix.tolist(), iy.tolist())), split=[], first=[],
               how=[])
    if len(ix) == 0:
        return res
    step = prm[4]
    cxy = [(prm[0] + x * step, prm[1] + y * step) for x, y in zip(ix, iy)]
    pg = planes_g[:6, iy, ix].T.astype(np.float32)
    po = planes_o[:6, iy, ix].T.astype(np.float32)
    ref_f = np.ascontiguousarray(ref, np.float32)
    dfm_f = np.ascontiguousarray(dfm, np.float32)

    def one_step(c, p):   # NO_SIMPLEX, the huge stop rule takes the first step as converged
        io, fo = oracle_solve_nodes(ref_f, dfm_f, dim, c, p, [2] * len(c), lm=lm, tol=1e30)
        return (io[:, 2] == 1), fo[:, :6]

    ok_g, st_g = one_step(cxy, pg)
    ok_o, st_o = one_step(cxy, po)
    g_first = ok_g & (np.abs(st_g[:, :2] - po[:, :2]).max(1) <= agree)
    o_first = ok_o & (np.abs(st_o[:, :2] - pg[:, :2]).max(1) <= agree)
    how = np.where(g_first, "gpu", np.where(o_first, "oracle", "-")).astype(object)
    if guess_g is not None:
        todo = [k for k in range(len(ix)) if how[k] == "-" and path_g[iy[k], ix[k]] == 3]
        if todo:
            c2 = [cxy[k] for k in todo]
            gs = guess_g[:, iy[todo], ix[todo]].T.astype(np.float32)
            md = [2 if tag_g[iy[k], ix[k]] in (2, 4) else 1 for k in todo]
            io_r, fo_r = oracle_solve_nodes(ref_f, dfm_f, dim, c2, gs, md, lm=lm, tol=tol)
            rr = fo_r[:, :6].astype(np.float32)
            near = np.abs(rr[:, :2] - pg[todo, :2]).max(1) <= agree_guess
            ok_a, st_a = one_step(c2, rr)
            ok_b, st_b = one_step(c2, pg[todo])
            a_b = ok_a & (np.abs(st_a[:, :2] - pg[todo, :2]).max(1) <= agree_guess)
            b_a = ok_b & (np.abs(st_b[:, :2] - rr[:, :2]).max(1) <= agree_guess)
            for j, k in enumerate(todo):
                if io_r[j, 1] == 0 and (near[j] or a_b[j] or b_a[j]):
                    how[k] = "guess"
    rest = [k for k in range(len(ix)) if how[k] == "-"]
    if rest:
        io_c, _ = oracle_solve_nodes(ref_f, dfm_f, dim, [cxy[k] for k in rest], pg[rest], [2] * len(rest), lm=lm,
                                     tol=tol)
        for j, k in enumerate(rest):
            if io_c[j, 1] == 0 and io_c[j, 2] == 1:
                how[k] = "converged"
    res["how"] = [str(h) for h in how]
    res["first"] = [h if h in ("gpu", "oracle") else "-" for h in res["how"]]
    res["split"] = [h in ("gpu", "oracle") for h in res["how"]]
    return res


def oracle_global_motion(ref, dfm, rect):
    """The oracle's global-motion model on a u8 pair: dict of locked, full,
    a (6: u = a0 + a1 dx + a2 dy, v = a3 + a4 dx + a5 dy about the rect
    centre), cx, cy, patches."""
    h, w = ref.shape
    out = np.zeros(12, np.float64)
    oracle().orc_global_motion(P(np.ascontiguousarray(ref, np.uint8)), P(np.ascontiguousarray(dfm, np.uint8)), w, h,
                               P(np.array(rect, np.int32)), P(out))
    return dict(locked=bool(out[0]), full=bool(out[1]), a=out[2:8].copy(), cx=out[8], cy=out[9],
                patches=int(out[10]))


def model_at(m, x, y):
    """(u, v) of a global-motion model at reference points (x, y)."""
    a = m["a"]
    dx, dy = np.asarray(x, np.float64) - m["cx"], np.asarray(y, np.float64) - m["cy"]
    return a[0] + a[1] * dx + a[2] * dy, a[3] + a[4] * dx + a[5] * dy


def oracle_vsg(planes, step, window):
    """The oracle's median test and VSG strain (pack_full_field_output's
    order) over a node field (7, gh, gw) of planes: (valid after the median
    test, exx, eyy, exy) on the grid."""
    gh, gw = planes.shape[1:]
    u = np.ascontiguousarray(planes[0], np.float32)
    v = np.ascontiguousarray(planes[1], np.float32)
    ok = np.ascontiguousarray(planes[6] >= 0, np.uint8)
    exx, eyy, exy = (np.zeros((gh, gw), np.float32) for _ in range(3))
    oracle().orc_outliers_and_vsg(gw, gh, step, window, P(u), P(v), P(ok), P(exx), P(eyy), P(exy))
    return ok.astype(bool), exx, eyy, exy


def fp64_window_errors(ref, dfm, prm, pts, planes, k=1500, seed=5, use6=False):
    """Each output point's error against the FP64 IC-GN optimum, on a seeded
    sample of k points: u and v against the point's own subset polished in
    FP64 (to ||dp|| < 1e-10 from its result), and the strains against the
    VSG strain (the oracle's median test and VSG) of the field whose every
    node in the point's strain window is polished. pts: (n, 8) output points;
    planes: (7, gh, gw) node planes of the same solve. Returns a dict of
    arrays (du, dv, dexx, deyy, dexy) over the sample."""
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(len(pts), size=min(k, len(pts)), replace=False))
    step, win = int(prm[4]), int(prm[6])
    gh, gw = planes.shape[1:]
    ix = ((pts[sel, 0] - prm[0]) / step).astype(int)
    iy = ((pts[sel, 1] - prm[1]) / step).astype(int)
    r = win / 2.0 / step
    rr = int(np.ceil(r))
    need = set()
    for x0, y0 in zip(ix, iy):
        for dy in range(-rr, rr + 1):
            for dx in range(-rr, rr + 1):
                x, y = x0 + dx, y0 + dy
                if 0 <= x < gw and 0 <= y < gh and planes[6, y, x] >= 0 and dx * dx + dy * dy <= r * r + 1e-9:
                    need.add((x, y))
    need = sorted(need)
    nx = np.array([x for x, _ in need])
    ny = np.array([y for _, y in need])
    si, p64, _ = fp64_polish(ref, dfm, int(prm[5]), [(prm[0] + x * step, prm[1] + y * step) for x, y in need],
                             planes[:6, ny, nx].T.astype(np.float64), use6=use6)
    ok = si[:, 0] == 0
    pf = planes.copy()
    pf[:6, ny[ok], nx[ok]] = p64[ok].T.astype(np.float32)
    _, ex, ey, exy = oracle_vsg(pf, step, win)
    return dict(du=np.abs(pts[sel, 2] - pf[0, iy, ix]), dv=np.abs(pts[sel, 3] - pf[1, iy, ix]),
                dexx=np.abs(pts[sel, 4] - ex[iy, ix]), deyy=np.abs(pts[sel, 5] - ey[iy, ix]),
                dexy=np.abs(pts[sel, 6] - exy[iy, ix]), polished=int(ok.sum()), needed=len(need))


def oracle_solve_nodes(ref, dfm, dim, nodes, guesses, modes, use6=False, lm=False, tol=1e-3):
    h, w = ref.shape
    n = len(nodes)
    cxy = np.array(nodes, np.int32).reshape(-1)
    g = np.array(guesses, np.float32).reshape(-1)
    md = np.array(modes, np.int32)
    io = np.zeros((n, 5), np.int32)
    fo = np.zeros((n, 7), np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    d = np.ascontiguousarray(dfm, np.float32)
    oracle().orc_solve_nodes(P(r), P(d), w, h, dim, int(use6), int(lm), ctypes.c_float(tol), n, P(cxy), P(g),
                             P(md), P(io), P(fo))
    return io, fo


def fp64_polish(ref, dfm, dim, nodes, p0, use6=False, tol=1e-10, max_iter=200):
    """FP64 IC-GN of the same model from per-node starts (the error yardstick)."""
    h, w = ref.shape
    n = len(nodes)
    cxy = np.array(nodes, np.int32).reshape(-1)
    start = np.ascontiguousarray(np.array(p0, np.float64).reshape(-1))
    si = np.zeros((n, 2), np.int32)
    po = np.zeros((n, 6), np.float64)
    corr = np.zeros(n, np.float64)
    r = np.ascontiguousarray(ref, np.float32)
    d = np.ascontiguousarray(dfm, np.float32)
    oracle().orc_icgn_fp64(P(r), P(d), w, h, dim, int(use6), n, P(cxy), P(start), ctypes.c_double(tol), max_iter,
                           P(si), P(po), P(corr))
    return si, po, corr


# ---- frame sequences with exact ground truth (tests/framework/sequence.h) ----

def stretch_rotate_aff(e, nu, theta_deg, tx, ty, cx, cy):
    """x = c + F (X - c) + t with F = R(theta) diag(1 + e, 1 - nu e), as the
    (u, v, ux, uy, vx, vy, cx, cy) affine the speckle renderer takes."""
    th = np.radians(theta_deg)
    c, s_ = np.cos(th), np.sin(th)
    a, d = 1.0 + e, 1.0 - nu * e
    F = np.array([[c * a, -s_ * d], [s_ * a, c * d]])
    return [tx, ty, F[0, 0] - 1.0, F[0, 1], F[1, 0], F[1, 1] - 1.0, cx, cy]


def aff_truth(aff, x, y):
    """Displacement and Green-Lagrange strain of a reference point under aff."""
    u0, v0, ux, uy, vx, vy, cx, cy = aff
    dx, dy = np.asarray(x, np.float64) - cx, np.asarray(y, np.float64) - cy
    u = u0 + ux * dx + uy * dy
    v = v0 + vx * dx + vy * dy
    exx = 0.5 * (2 * ux + ux * ux + vx * vx)
    eyy = 0.5 * (2 * vy + uy * uy + vy * vy)
    exy = 0.5 * (uy + vx + ux * uy + vx * vy)
    return u, v, exx, eyy, exy


def speckle_sequence(T, w=256, h=256, e_max=0.1, theta_max=4.0, nu=0.3, tx_max=0.0, ty_max=0.0, seed=31,
                     blobs=2400):
    """Frames 1..T of stretch e_max t / T with rotation theta_max t / T and a
    drift (tx, ty) t / T about the image centre, rendered from one speckle."""
    ref = gray8(ref_speckle(seed, w, h, blobs))
    frames, affs = [], []
    for t in range(1, T + 1):
        f = t / T
        aff = stretch_rotate_aff(e_max * f, nu, theta_max * f, tx_max * f, ty_max * f, w / 2.0, h / 2.0)
        frames.append(gray8(ref_speckle(seed, w, h, blobs, aff)))
        affs.append(aff)
    return ref, frames, affs


def centred_params(w, h, roi, step=8, subset=21, window=None, use6=False):
    return ((w - roi) // 2, (h - roi) // 2, roi, roi, step, subset, window or 2 * step + 1, 1 if use6 else 0)


class OracleIncremental:
    """The CPU restatement of IncrementalSession (oracle/src/incremental.cpp)."""

    def __init__(self):
        self.lib = oracle()
        self.h = self.lib.orc_inc_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_inc_destroy(self.h)

    def init(self, ref, prm, interval, mask=None, threads=0, tol=1e-3):
        ref = np.ascontiguousarray(ref, np.uint8)
        h, w = ref.shape
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        pr = np.array(prm, np.int32)
        self.cap = (prm[2] // prm[4] + 1) * (prm[3] // prm[4] + 1) * 8
        return self.lib.orc_inc_init(self.h, P(ref), w, h, P(m), P(pr), interval, threads, ctypes.c_float(tol))

    def process(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros(self.cap, np.float32)
        met = np.zeros(24, np.float32)
        info = np.zeros(6, np.int32)
        n = self.lib.orc_inc_process(self.h, P(img), P(out), self.cap, P(met), 24, P(info))
        return n, out[: max(n, 0) * 8].reshape(-1, 8).copy(), met, info


# ---- DICe open-hole-tension fixtures (tests/fixtures/dice, LICENSE.DICe) ----

DICE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dice")
# (image, DICe solution) per load step, each against oht_cfrp_00.tiff
DICE_LADDER = [("oht_cfrp_01.tiff", "DICe_solution_01.txt"), ("oht_cfrp_03.tiff", "DICe_solution_03.txt"),
               ("oht_cfrp_06.tiff", "DICe_solution_06.txt"), ("oht_cfrp_11.tiff", "DICe_solution_11.txt")]


def dice_bytes(name):
    with open(os.path.join(DICE_DIR, name), "rb") as f:
        return f.read()


def dice_image(name):
    """A fixture decoded by PIL (independent of the engine's codec)."""
    import io

    from PIL import Image

    return np.array(Image.open(io.BytesIO(dice_bytes(name))).convert("L"))


def dice_gold(name):
    """DICe's solved field: rows (x, y, u, v); the '***' header and the column
    line skipped, x and y rounded (test_field_agreement.cpp:63-81)."""
    rows = []
    with open(os.path.join(DICE_DIR, name)) as f:
        for line in f:
            if not line.strip() or line.startswith("***") or line.startswith("COORDINATE_X"):
                continue
            c = line.split(",")
            if len(c) < 4:
                continue
            rows.append((round(float(c[0])), round(float(c[1])), float(c[2]), float(c[3])))
    return np.array(rows)


def dice_grid_params():
    """The step-7 grid whose nodes include every DICe point two steps inside
    the ROI (test_incremental_oht.cpp:86-98)."""
    step = 7
    return (61 - 2 * step, 61 - 2 * step, (341 - 61) + 4 * step, (971 - 61) + 4 * step, step, 27, 21, 1)


def dice_agreement(pts, gold):
    """(compared fraction, rms du, rms dv, mean du, max |d|) of an (n, 8) point
    list at DICe's coordinates (test_incremental_oht.cpp:104-135)."""
    at = {(int(round(r[0])), int(round(r[1]))): (float(r[2]), float(r[3])) for r in pts}
    d = np.array([(at[(int(x), int(y))][0] - u, at[(int(x), int(y))][1] - v)
                  for x, y, u, v in gold if (int(x), int(y)) in at])
    if len(d) == 0:
        return 0.0, np.inf, np.inf, np.inf, np.inf
    rms = np.sqrt((d ** 2).mean(0))
    return len(d) / len(gold), float(rms[0]), float(rms[1]), float(d[:, 0].mean()), float(np.abs(d).max())


# ---- the learned seeder's CPU restatement (oracle/src/seeder.cpp) ---------------

def seeder_params(seed=0):
    """He-normal conv weights and a 1/sqrt(512) linear layer, small random
    biases, in the flat layout of csrc/seeder.hpp."""
    lib = oracle()
    lib.orc_seeder_param_count.restype = ctypes.c_int
    n = lib.orc_seeder_param_count()
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    o = 0
    for l in range(4):
        cin = 2 if l == 0 else 32
        nw = 32 * cin * 9
        p[o:o + nw] = rng.normal(0, np.sqrt(2.0 / (cin * 9)), nw)
        o += nw
        p[o:o + 32] = rng.normal(0, 0.05, 32)
        o += 32
    p[o:o + 6 * 512] = rng.normal(0, np.sqrt(1.0 / 512), 6 * 512)
    o += 6 * 512
    p[o:o + 6] = rng.normal(0, 0.05, 6)
    return p


def seeder_batch(B, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (B, 2, 64, 64)).astype(np.float32)
    t = rng.normal(0, 1, (B, 6)).astype(np.float32)
    return x, t


def oracle_seeder(params, x, t=None):
    """(loss, y, grads) of the double-precision restatement."""
    lib = oracle()
    lib.orc_seeder_pass.restype = ctypes.c_double
    B = x.shape[0]
    y = np.zeros((B, 6), np.float32)
    g = np.zeros_like(params)
    loss = lib.orc_seeder_pass(P(params), P(np.ascontiguousarray(x)), P(t) if t is not None else None, B, P(y),
                               P(g) if t is not None else None)
    return loss, y, g
