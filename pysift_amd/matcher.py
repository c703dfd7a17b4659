"""Exact mutual ratio-test matcher on uint8 descriptors (INT8 MFMA).

Restates GPUPyStitch._match_features / _match_symmetric (reference
pysift/core.py:2840-2940, 2652-2701) on the exact integer distance
d = sum (a - b)^2 of uint8 descriptors: ratio test d1 < ratio^2 d2 in both
directions plus the mutual check, matches ordered by query index.
"""
import ctypes

from . import _lib

METHODS = {"hilo": 0, "shift": 1, "twopass": 2}
DEFAULT_METHOD = "twopass"
# two-pass tile shapes (0: 64 rows a wave, next tile's MFMAs issued ahead;
# 1: 64 rows; 2: 32 rows; 3: 96 rows; 4: 128 rows) and 512-column chunks per
# workgroup tried by autotune
TILE_SHAPES = (0, 1, 2, 3, 4)
TILE_CHUNKS = (1, 2, 4, 8)


def match_workspace_bytes(N, M):
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().msift_match_workspace_bytes(int(N), int(M), ctypes.byref(n)), "match workspace")
    return n.value


def match_u8(a, b, ratio=0.75, method=DEFAULT_METHOD, sync=True, ws=None):
    """a: (N,128) uint8 CUDA tensor, b: (M,128).  Returns (fwd (N,3), bwd (M,3),
    pairs (K,2)) int32 device tensors; fwd[i] = (best j, d1, d2)."""
    import torch
    if a.dtype != torch.uint8 or b.dtype != torch.uint8 or a.dim() != 2 or b.dim() != 2 \
            or a.shape[1] != 128 or b.shape[1] != 128:
        raise ValueError("descriptors must be (N,128) uint8")
    a = a.contiguous(); b = b.contiguous()
    dev = a.device
    N, M = a.shape[0], b.shape[0]
    if ws is None or ws.numel() < match_workspace_bytes(N, M):
        ws = torch.empty(max(match_workspace_bytes(N, M), 1), dtype=torch.uint8, device=dev)
    fwd = torch.empty((max(N, 1), 3), dtype=torch.int32, device=dev)
    bwd = torch.empty((max(M, 1), 3), dtype=torch.int32, device=dev)
    pairs = torch.empty((max(min(N, M), 1), 2), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = _lib.lib().msift_match_u8(_lib.ptr(a), N, _lib.ptr(b), M, float(ratio), METHODS[method],
                                   _lib.ptr(ws), ws.numel(), _lib.ptr(fwd), _lib.ptr(bwd), _lib.ptr(pairs),
                                   _lib.ptr(npairs), _lib.stream_ptr(dev))
    _lib.check(rc, "match_u8")
    if not sync:
        return fwd[:N], bwd[:M], pairs, npairs
    k = int(npairs.item())
    return fwd[:N], bwd[:M], pairs[:k]


def set_tile(shape, chunks_per_block):
    """Two-pass tile shape for the process (outputs do not depend on it)."""
    _lib.check(_lib.lib().msift_match_set_tile(int(shape), int(chunks_per_block)), "match set_tile")


def _bucket(n):
    b = 1024
    while b < n:
        b *= 2
    return b


def autotune(N, M, device="cuda", iters=3):
    """Time the two-pass matcher at N x M for every tile shape and chunk count,
    keep the fastest (cached per power-of-two size class in
    ~/.cache/pysift_amd/autotune.json) and set it.  Returns (shape, chunks),
    timings in ms.  The kernels' time does not depend on the descriptor
    values, so random descriptors of the requested size are timed."""
    import torch
    from .batch import _load_tune, _save_tune
    key = f"match5:{_bucket(N)}:{_bucket(M)}"
    cache = _load_tune()
    if key in cache:
        shape, cpw = cache[key]
        set_tile(shape, cpw)
        return (shape, cpw), {}
    g = torch.Generator().manual_seed(0)
    a = torch.randint(0, 256, (max(N, 1), 128), dtype=torch.uint8, generator=g).to(device)
    b = torch.randint(0, 256, (max(M, 1), 128), dtype=torch.uint8, generator=g).to(device)
    ws = torch.empty(match_workspace_bytes(a.shape[0], b.shape[0]), dtype=torch.uint8, device=device)
    times = {}
    with torch.cuda.device(a.device):
        for shape in TILE_SHAPES:
            for cpw in TILE_CHUNKS:
                set_tile(shape, cpw)
                match_u8(a, b, 0.75, "twopass", sync=False, ws=ws)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(iters):
                    match_u8(a, b, 0.75, "twopass", sync=False, ws=ws)
                e1.record()
                torch.cuda.synchronize()
                times[(shape, cpw)] = e0.elapsed_time(e1) / iters
    best = min(times, key=times.get)
    set_tile(*best)
    cache[key] = list(best)
    _save_tune(cache)
    return best, times


def match_f16(a, b, ratio=0.75, ws=None):
    """Float descriptors a: (N,128), b: (M,128) float32 CUDA tensors (unit
    norm, e.g. RootSIFT) matched on FP16 MFMA with FP32 accumulation.
    Returns (fwd (N,3), bwd (M,3), pairs (K,2)); the distance columns are the
    float32 squared distances viewed as int32 (use .view(torch.float32))."""
    import torch
    if a.dtype != torch.float32 or b.dtype != torch.float32 or a.dim() != 2 or b.dim() != 2 \
            or a.shape[1] != 128 or b.shape[1] != 128:
        raise ValueError("descriptors must be (N,128) float32")
    a = a.contiguous(); b = b.contiguous()
    dev = a.device
    N, M = a.shape[0], b.shape[0]
    need = match_workspace_bytes(N, M)
    if ws is None or ws.numel() < need:
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
    fwd = torch.empty((max(N, 1), 3), dtype=torch.int32, device=dev)
    bwd = torch.empty((max(M, 1), 3), dtype=torch.int32, device=dev)
    pairs = torch.empty((max(min(N, M), 1), 2), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = _lib.lib().msift_match_f16(_lib.ptr(a), N, _lib.ptr(b), M, float(ratio), _lib.ptr(ws), ws.numel(),
                                    _lib.ptr(fwd), _lib.ptr(bwd), _lib.ptr(pairs), _lib.ptr(npairs),
                                    _lib.stream_ptr(dev))
    _lib.check(rc, "match_f16")
    k = int(npairs.item())
    return fwd[:N], bwd[:M], pairs[:k]


class MatchBatchPlan:
    """Plan of a batched match over pairs (a_p: N[p] x 128, b_p: M[p] x 128
    uint8): one launch over every tile of every pair; per pair the results
    equal match_u8 on that pair."""

    def __init__(self, Ns, Ms, device=None):
        import torch
        self.Ns = [int(n) for n in Ns]
        self.Ms = [int(m) for m in Ms]
        if len(self.Ns) != len(self.Ms) or not self.Ns:
            raise ValueError("need equal, non-empty lists of pair sizes")
        self.np = len(self.Ns)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        L = _lib.lib()
        arrN = (ctypes.c_int * self.np)(*self.Ns)
        arrM = (ctypes.c_int * self.np)(*self.Ms)
        p = ctypes.c_void_p()
        _lib.check(L.msift_match_batch_create(self.np, arrN, arrM, ctypes.byref(p)), "match batch")
        self._p = p
        ws = ctypes.c_size_t()
        ao = (ctypes.c_int * self.np)()
        bo = (ctypes.c_int * self.np)()
        _lib.check(L.msift_match_batch_sizes(p, ctypes.byref(ws), ao, bo), "match batch sizes")
        self.a_off = list(ao)
        self.b_off = list(bo)
        dev = self.device
        sN, sM = sum(self.Ns), sum(self.Ms)
        self.ws = torch.empty(max(ws.value, 1), dtype=torch.uint8, device=dev)
        self.fwd = torch.empty((max(sN, 1), 3), dtype=torch.int32, device=dev)
        self.bwd = torch.empty((max(sM, 1), 3), dtype=torch.int32, device=dev)
        self.pairs = torch.empty((max(sN, 1), 2), dtype=torch.int32, device=dev)
        self.npairs = torch.zeros(self.np, dtype=torch.int32, device=dev)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and _lib._lib is not None:
            _lib._lib.msift_match_batch_destroy(p)
            self._p = None

    def pack(self, sets, counts):
        """Concatenate (n_p, 128) uint8 tensors into one (sum n_p, 128) tensor."""
        import torch
        for t, n in zip(sets, counts):
            if t.dtype != torch.uint8 or t.dim() != 2 or t.shape[1] != 128 or t.shape[0] != n:
                raise ValueError("descriptor sets must be (n,128) uint8 matching the plan")
        if sum(counts) == 0:
            return torch.zeros((1, 128), dtype=torch.uint8, device=self.device)
        return torch.cat([t.to(self.device) for t in sets]).contiguous()

    def run(self, A, B, ratio=0.75, method=DEFAULT_METHOD):
        """A, B packed (sum N, 128) / (sum M, 128) uint8 device tensors."""
        rc = _lib.lib().msift_match_u8_batched(
            self._p, _lib.ptr(A), _lib.ptr(B), float(ratio), METHODS[method], _lib.ptr(self.ws), self.ws.numel(),
            _lib.ptr(self.fwd), _lib.ptr(self.bwd), _lib.ptr(self.pairs), _lib.ptr(self.npairs),
            _lib.stream_ptr(self.device))
        _lib.check(rc, "match_u8_batched")

    def results(self):
        """Per pair (fwd (N,3), bwd (M,3), pairs (K,2)) device tensors (synchronises)."""
        k = self.npairs.cpu().tolist()
        out = []
        for p in range(self.np):
            a, b, n, m = self.a_off[p], self.b_off[p], self.Ns[p], self.Ms[p]
            out.append((self.fwd[a:a + n], self.bwd[b:b + m], self.pairs[a:a + k[p]]))
        return out

    def gather_points(self, kpa, kpb):
        """Keypoint rows (x, y, ...) packed like the descriptors -> (p1, p2)
        (sum N, 2) float64 with pair p's matches at rows a_off[p] ..."""
        import torch
        kpa = kpa.contiguous()
        kpb = kpb.contiguous()
        sN = max(sum(self.Ns), 1)
        p1 = torch.zeros((sN, 2), dtype=torch.float64, device=self.device)
        p2 = torch.zeros((sN, 2), dtype=torch.float64, device=self.device)
        rc = _lib.lib().msift_match_batch_gather_points(
            self._p, _lib.ptr(kpa), kpa.shape[1], _lib.ptr(kpb), kpb.shape[1], _lib.ptr(self.pairs),
            _lib.ptr(self.npairs), _lib.ptr(p1), _lib.ptr(p2), _lib.stream_ptr(self.device))
        _lib.check(rc, "gather points")
        return p1, p2


def match_u8_batched(sets_a, sets_b, ratio=0.75, method=DEFAULT_METHOD):
    """Match pairs (sets_a[p], sets_b[p]) in one launch; returns per pair
    (fwd, bwd, pairs) as match_u8 would."""
    plan = MatchBatchPlan([t.shape[0] for t in sets_a], [t.shape[0] for t in sets_b], device=sets_a[0].device)
    A = plan.pack(sets_a, plan.Ns)
    B = plan.pack(sets_b, plan.Ms)
    plan.run(A, B, ratio, method)
    return plan.results()


# ---- the reference's float matcher (core.py:2840-2940, int8_matching off) ----

def l2_normalize(x):
    """(n, 128) float32 CUDA tensor -> x / (||x|| + 1e-7) per row (core.py:2879-2880)."""
    import torch
    if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != 128:
        raise ValueError("descriptors must be (n,128) float32")
    x = x.contiguous()
    y = torch.empty_like(x)
    _lib.check(_lib.lib().msift_l2_normalize(_lib.ptr(x), x.shape[0], _lib.ptr(y), _lib.stream_ptr(x.device)),
               "l2_normalize")
    return y


def pca_fit(a, b, dims=64):
    """_pca_compress's fit (core.py:2728-2736) on the L2-normalised rows of
    a (N,128) and b (M,128) float32 CUDA tensors stacked: FP64 mean and
    covariance, Jacobi eigenvectors.  Returns (mean (128,) f32, proj (128,
    dims) f32, eigenvalues (128,) f64 descending, sweeps (1,) int32), all on
    the device."""
    import torch
    for t in (a, b):
        if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 128:
            raise ValueError("descriptors must be (n,128) float32")
    a = a.contiguous(); b = b.contiguous()
    dev = a.device
    N, M = a.shape[0], b.shape[0]
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().msift_pca_workspace_bytes(N, M, ctypes.byref(n)), "pca workspace")
    ws = torch.empty(n.value, dtype=torch.uint8, device=dev)
    mean = torch.empty(128, dtype=torch.float32, device=dev)
    proj = torch.empty((128, int(dims)), dtype=torch.float32, device=dev)
    evals = torch.empty(128, dtype=torch.float64, device=dev)
    sweeps = torch.empty(1, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().msift_pca_fit(_lib.ptr(a), N, _lib.ptr(b), M, int(dims), _lib.ptr(ws), ws.numel(),
                                        _lib.ptr(mean), _lib.ptr(proj), _lib.ptr(evals), _lib.ptr(sweeps),
                                        _lib.stream_ptr(dev)), "pca_fit")
    return mean, proj, evals, sweeps


def pca_project(x, mean, proj):
    """Unit rows (n,128) -> ((x - mean) @ proj) re-normalised with +1e-7 (core.py:2737-2742)."""
    import torch
    x = x.contiguous()
    dims = proj.shape[1]
    y = torch.empty((x.shape[0], dims), dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().msift_pca_project(_lib.ptr(x), x.shape[0], _lib.ptr(mean), _lib.ptr(proj), dims,
                                            _lib.ptr(y), _lib.stream_ptr(x.device)), "pca_project")
    return y


def match_sim_f16(a, b, ratio=0.75, ws=None):
    """_match_symmetric (core.py:2652-2701) on unit rows a (N,D), b (M,D)
    float32, D <= 128: binary16 similarity on FP16 MFMA, top-2 by similarity
    both ways, ratio test on d = max(2 - 2 s, 0), mutual pairs.  Returns
    (fwd (N,3), bwd (M,3), pairs (K,2)); the distance columns hold d as
    float32 bits (view them with .view(torch.float32))."""
    import torch
    if a.dtype != torch.float32 or b.dtype != torch.float32 or a.dim() != 2 or b.dim() != 2 \
            or a.shape[1] != b.shape[1] or not 1 <= a.shape[1] <= 128:
        raise ValueError("descriptors must be (N,D) and (M,D) float32 with D <= 128")
    a = a.contiguous(); b = b.contiguous()
    dev = a.device
    N, M, D = a.shape[0], b.shape[0], a.shape[1]
    need = match_workspace_bytes(N, M)
    if ws is None or ws.numel() < need:
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
    fwd = torch.empty((max(N, 1), 3), dtype=torch.int32, device=dev)
    bwd = torch.empty((max(M, 1), 3), dtype=torch.int32, device=dev)
    pairs = torch.empty((max(min(N, M), 1), 2), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().msift_match_sim_f16(_lib.ptr(a), N, _lib.ptr(b), M, D, float(ratio), _lib.ptr(ws),
                                              ws.numel(), _lib.ptr(fwd), _lib.ptr(bwd), _lib.ptr(pairs),
                                              _lib.ptr(npairs), _lib.stream_ptr(dev)), "match_sim_f16")
    k = int(npairs.item())
    return fwd[:N], bwd[:M], pairs[:k]


def match_float(da, db, ratio=0.75, pca_dims=64, pca=None):
    """The reference's float matching path end to end (core.py:2875-2939 with
    int8_matching off): normalise, PCA to pca_dims (0 or >= 128: none; fitted
    on this pair unless pca = (mean, proj) is given, as the reference caches
    _pca_matrix after the first pair), binary16 similarity, symmetric ratio
    test, mutual pairs.  Returns (fwd, bwd, pairs, pca)."""
    a, b = l2_normalize(da), l2_normalize(db)
    if 0 < pca_dims < 128:
        if pca is None:
            mean, proj, _, _ = pca_fit(da, db, pca_dims)
            pca = (mean, proj)
        a, b = pca_project(a, *pca), pca_project(b, *pca)
    fwd, bwd, pairs = match_sim_f16(a, b, ratio)
    return fwd, bwd, pairs, pca
