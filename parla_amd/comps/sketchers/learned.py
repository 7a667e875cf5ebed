"""Learned leverage-score row sampling with an SJLT fallback.

No counterpart in the reference.  The reference's data-oblivious row sampler
(SkOpIN, comps/sketchers/oblivious.py:75-82, utils/sketching.py:204-232) draws
rows uniformly; this module draws them from leverage scores predicted by a
small MLP and keeps the reference's plain SJLT sketch as the fallback:

  1. features: F = A G (G a 16-column Gaussian pre-sketch, one MFMA GEMM over A),
     normalised by their root-mean-square;
  2. the MLP 16 -> 64 -> 64 -> 1 on asinh of the features (BF16 weights, FP32
     arithmetic, the library's own forward/backward kernels) predicts
     log(leverage * m / n) per row, capped at log(m / n) (leverage <= 1);
  3. d rows are sampled with replacement from p = (1 - mix) softmax(y) + mix/m
     and scaled by 1/sqrt(d p_i), so E[S^T S] = I (the importance-sampling
     sketch of utils/sketching.py:204-232 with weights);
  4. SPOLearned factors S A, estimates kappa(A R^{-1}) by a few Lanczos (CG)
     steps on R^{-T} A^T A R^{-1}, and falls back to the plain SJLT SPO when
     the estimate exceeds `kappa_max`.

The network is trained on exact leverage scores ||Q_i||^2 from the library's
FP64 QR (CholeskyQR2 / Householder) of matrices of the target family.
"""
import time

import numpy as np
import torch

from parla_amd import ops
from parla_amd.comps.determiter.logging import SketchAndPrecondLog
import parla_amd.comps.determiter.saddle as dsad

MLP_IN, MLP_H = 16, 64
_SIZES = [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in _SIZES)


def _bf16_bits(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class LeverageMLP:
    """The surrogate's parameters on the device: FP32 masters and Adam moments,
    BF16 weights and FP32 biases for the kernels."""

    def __init__(self, seed=0, dev=None):
        self.dev = dev or ops.default_device()
        rng = np.random.default_rng(seed)
        parts = []
        for name, sh in _SIZES:
            if name == "W1":
                parts.append(rng.normal(0, np.sqrt(2.0 / 16), sh))
            elif name == "W2":
                parts.append(rng.normal(0, np.sqrt(2.0 / 64), sh))
            elif name == "w3":
                parts.append(rng.normal(0, np.sqrt(1.0 / 64), sh))
            else:
                parts.append(np.zeros(sh))
        self.set_theta(np.concatenate([p.reshape(-1) for p in parts]))

    def set_theta(self, theta):
        theta = np.asarray(theta, dtype=np.float32)
        if theta.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters")
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.w = torch.tensor(theta, **f32)
        self.m = torch.zeros(NPARAM, **f32)
        self.v = torch.zeros(NPARAM, **f32)
        self.step = 0
        o = 0
        wparts, bparts = [], []
        for name, sh in _SIZES:
            k = int(np.prod(sh))
            (wparts if name in ("W1", "W2", "w3") else bparts).append(theta[o:o + k])
            o += k
        self.wb = torch.from_numpy(_bf16_bits(np.concatenate(wparts)).view(np.int16)).to(self.dev)
        self.bias = torch.tensor(np.concatenate(bparts), **f32)

    def theta(self):
        return self.w.cpu().numpy()

    def predict(self, F):
        rows = F.shape[0]
        y = torch.empty(rows, dtype=torch.float32, device=self.dev)
        ops.check(ops.lib().pamd_mlp_fwd(rows, ops._p(F), F.stride(0), ops._p(self.wb), ops._p(self.bias),
                                         ops._p(y), ops.stream(self.dev)), "pamd_mlp_fwd")
        return y

    def loss_grad(self, F, t, rows_total=None):
        """(gradient (NPARAM,), loss (device scalar)) of (1/N) sum (MLP(F) - t)^2."""
        rows = F.shape[0]
        ws = ops.workspace(ops._query(ops.lib().pamd_mlp_bwd_workspace, rows), self.dev)
        g = torch.empty(NPARAM + 1, dtype=torch.float32, device=self.dev)
        ops.check(ops.lib().pamd_mlp_bwd(rows, rows_total or rows, ops._p(F), F.stride(0), ops._p(t),
                                         ops._p(self.wb), ops._p(self.bias), ops._p(g), ops._p(ws),
                                         ops.stream(self.dev)), "pamd_mlp_bwd")
        return g[:NPARAM], g[NPARAM]

    def adam(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        ops.check(ops.lib().pamd_mlp_adam(ops._p(self.w), ops._p(self.m), ops._p(self.v), ops._p(grad), lr, beta1,
                                          beta2, eps, self.step, ops._p(self.wb), ops._p(self.bias),
                                          ops.stream(self.dev)), "pamd_mlp_adam")

    def fit(self, data, epochs=4, lr=2e-3, batch_rows=65536, seed=0, log=None):
        """Adam over the rows of [(F, t), ...] (device FP32 features and labels)."""
        rng = np.random.default_rng(seed)
        chunks = [(i, r0) for i, (F, _) in enumerate(data) for r0 in range(0, F.shape[0], batch_rows)]
        losses = []
        for ep in range(epochs):
            order = rng.permutation(len(chunks))
            tot = torch.zeros((), dtype=torch.float32, device=self.dev)
            for ci in order:
                i, r0 = chunks[ci]
                F, t = data[i]
                Fb, tb = F[r0:r0 + batch_rows], t[r0:r0 + batch_rows]
                g, loss = self.loss_grad(Fb, tb)
                self.adam(g, lr=lr)
                tot += loss
            losses.append(float(tot.item()) / len(chunks))
            if log:
                log(f"epoch {ep}: mean batch loss {losses[-1]:.4f}")
        return losses


# ---------------------------------------------------------------------------
# features, labels and the training family
# ---------------------------------------------------------------------------

def presketch(n, seed, dev):
    """The 16-column Gaussian pre-sketch G (n x 16), N(0, 1/16) entries."""
    return ops.gaussian_fill(n, MLP_IN, 1.0 / 4.0, seed, dev)


def features(A, G):
    """F = A G normalised to unit root-mean-square, FP32 (m x 16)."""
    F = ops.gemm(ops.as_device(A), G)
    rms = torch.sqrt(torch.mean(F * F))
    return (F / rms).to(torch.float32).contiguous()


def leverage_labels(A, rounds=2):
    """t_i = log(leverage_i * m / n), leverage_i = ||Q_i||^2 from the FP64
    CholeskyQR of A (`rounds` rounds: after one round Q is orthonormal to
    kappa(A)^2 u, after two to u while kappa(A) << 1e8); a failed pivot falls
    back to the library's orth (CholeskyQR2 / Householder)."""
    m, n = A.shape
    Q = A
    info = torch.zeros(rounds, dtype=torch.int32, device=A.device)
    for r in range(rounds):
        G = ops.gemm(Q.t(), Q)
        ops.potrf_(G, info[r:r + 1])
        Q = ops.trsm_mfma(Q, G)
    if bool(info.any().item()):
        Q = ops.orth_q(A)
    lev = torch.sum(Q * Q, dim=1)
    return torch.log(torch.clamp(lev * (m / n), min=1e-30)).to(torch.float32).contiguous()


def family_matrix(m, n, seed, dev, heavy_frac=0.01, heavy_scale=100.0, cond=1e6):
    """A = D G diag(s) V^T: Gaussian rows G (device Philox), D scales a seeded
    1% of rows by 100, s log-spaced from 1 to 1/cond, V Haar; A is returned in
    FP32, the family's data type (the FP64 solvers widen it exactly).  Returns
    (A, D G): leverage scores do not depend on the right factor, so labels
    come from the better-conditioned D G."""
    rng = np.random.default_rng(seed)
    G = ops.gaussian_fill(m, n, 1.0, int(rng.integers(0, 2 ** 63)), dev)
    heavy = torch.from_numpy(rng.choice(m, max(1, int(heavy_frac * m)), replace=False)).to(dev)
    G[heavy] *= heavy_scale
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    M = torch.from_numpy(np.ascontiguousarray(np.logspace(0, -np.log10(cond), n)[:, None] * V.T)).to(dev)
    return ops.gemm(G, M).to(torch.float32), G


def training_set(num, m, n, seed, dev, log=None):
    data = []
    for i in range(num):
        A, DG = family_matrix(m, n, seed + i, dev)
        Gp = presketch(n, seed + 10_000_000 + i, dev)
        data.append((features(A, Gp), leverage_labels(DG, rounds=1)))  # kappa(D G) < 10: one round is exact
        del A, DG
        if log and (i + 1) % max(1, num // 10) == 0:
            log(f"training set: {i + 1}/{num} matrices")
    return data


# ---------------------------------------------------------------------------
# sampling sketch and the driver
# ---------------------------------------------------------------------------

def sampling_probabilities(y, mix=0.1, n=None):
    """p = (1 - mix) softmax(y) + mix / m (float64, on the device).  With n
    given, predictions are first capped at log(m / n): a leverage score never
    exceeds 1, and the cap stops a few over-predicted rows from taking the
    whole distribution."""
    y = y.to(torch.float64)
    if n is not None:
        y = torch.clamp(y, max=float(np.log(y.numel() / n)))
    p = torch.softmax(y, dim=0)
    return (1.0 - mix) * p + mix / y.numel()


def sample_rows(p, d, rng, backend="device"):
    """d row indices drawn with replacement from the weights p and their
    scales 1/sqrt(d p_i / sum p).  backend 'device': inverse CDF on Philox
    uniforms in the library (pamd_sample_rows_f64; seed from rng), results on
    the device, no host round trip; 'numpy': numpy's Generator.choice on the
    host, results as numpy arrays."""
    if backend == "numpy":
        ph = p.cpu().numpy()
        ph = ph / ph.sum()
        idx = rng.choice(ph.size, size=d, replace=True, p=ph)
        return idx, 1.0 / np.sqrt(d * ph[idx])
    dev = p.device
    pc = p.to(torch.float64).contiguous()
    m = pc.numel()
    idx = torch.empty(d, dtype=torch.int64, device=dev)
    scale = torch.empty(d, dtype=torch.float64, device=dev)
    ws = ops.workspace(ops._query(ops.lib().pamd_sample_rows_workspace, m), dev)
    ops.check(ops.lib().pamd_sample_rows_f64(m, ops._p(pc), d, int(rng.integers(0, 2 ** 63)), ops._p(idx),
                                             ops._p(scale), ops._p(ws), ops.stream(dev)), "pamd_sample_rows_f64")
    return idx, scale


def row_gather(A, idx, scale, out):
    dev = A.device
    it = idx if isinstance(idx, torch.Tensor) else torch.from_numpy(np.asarray(idx, dtype=np.int64)).to(dev)
    st = scale if isinstance(scale, torch.Tensor) else \
        torch.from_numpy(np.asarray(scale, dtype=np.float64)).to(dev)
    m, n = A.shape
    ops.check(ops.lib().pamd_row_gather_f64(len(idx), m, n, ops._p(it), ops._p(st), ops._p(A), A.stride(0),
                                            ops._p(out), out.stride(0), ops.stream(dev)), "pamd_row_gather_f64")
    return out


def kappa_estimate(A, R, steps=12, seed=0):
    """Estimate kappa(A R^{-1}) from the Lanczos tridiagonal that CG on
    R^{-T} A^T A R^{-1} builds (extreme Ritz values); steps passes over A.
    R^{-1} is formed once (the estimate needs only its leading digits), so a
    step is one normal pass and two products with it instead of two
    triangular solves."""
    n = R.shape[0]
    normal = ops.NormalOp(A)
    Ri = ops.trtri(R) if n <= 2048 else None

    def apply(v):
        if Ri is not None:
            return ops.gemv(Ri, normal(ops.gemv(Ri, v)), trans=True)
        return ops.trsv_(R, normal(ops.trsv_(R, v.clone())), trans=True)

    rng = np.random.default_rng(seed)
    r = torch.from_numpy(rng.standard_normal(n)).to(A.device)
    p = r.clone()
    rr = torch.dot(r, r)
    alphas, betas = [], []
    for _ in range(steps):
        q = apply(p)
        alpha = rr / torch.dot(p, q)
        r = r - alpha * q
        rr_new = torch.dot(r, r)
        beta = rr_new / rr
        p = r + beta * p
        rr = rr_new
        alphas.append(alpha)
        betas.append(beta)
    a = torch.stack(alphas).cpu().numpy()
    b = torch.stack(betas).cpu().numpy()
    k = len(a)
    T = np.zeros((k, k))
    for i in range(k):
        T[i, i] = 1.0 / a[i] + (b[i - 1] / a[i - 1] if i > 0 else 0.0)
        if i + 1 < k:
            T[i, i + 1] = T[i + 1, i] = np.sqrt(max(b[i], 0.0)) / a[i]
    ev = np.linalg.eigvalsh(T)
    return float(np.sqrt(ev[-1] / max(ev[0], 1e-300)))


class SPOLearned:
    """SPO (least_squares.py:206-369 call contract) with the learned row-sampling
    sketch and the plain SJLT SPO as the fallback."""

    def __init__(self, mlp, d=4096, kappa_max=10.0, mix=0.1, fallback=None, lanczos_steps=4, sampling="device"):
        from parla_amd.drivers.least_squares import SPO
        import parla_amd.comps.sketchers.oblivious as sko
        self.mlp = mlp
        self.d = d
        self.kappa_max = kappa_max
        self.mix = mix
        self.lanczos_steps = lanczos_steps
        self.sampling = sampling
        self.fallback = fallback or SPO(sko.SkOpSJ(8), 4)
        self.iterative_solver = dsad.PcSS2()

    def sketch_factor(self, At, bt, rng, probs=None):
        """Row-sampling sketch of [A | b] and its R and z = Q^T S b; returns (R, z, kappa_est)."""
        At = ops.as_device(At)
        m, n = At.shape
        dev = At.device
        if probs is None:
            G = presketch(n, int(rng.integers(0, 2 ** 63)), dev)
            probs = sampling_probabilities(self.mlp.predict(features(At, G)), self.mix, n)
        idx, scale = sample_rows(probs, self.d, rng, self.sampling)
        Mx = torch.empty((self.d, n + 1), dtype=ops.F64, device=dev)
        row_gather(At, idx, scale, Mx[:, :n])
        it = idx if isinstance(idx, torch.Tensor) else torch.from_numpy(idx).to(dev)
        st = scale if isinstance(scale, torch.Tensor) else torch.from_numpy(scale).to(dev)
        row_gather(bt.reshape(-1, 1), it, st, Mx[:, n:])
        # R and Q^T S b of the sample: CholeskyQR2 when the sample is well enough
        # conditioned for it (the learned or row-norm probabilities keep it
        # near kappa(A)), else Householder
        rz = ops.sketch_rz_cholqr2(Mx, n)
        if rz is not None:
            R, z = rz
        else:
            fac = ops.qr_r_(Mx, n)
            R = torch.triu(fac.R_view).contiguous()
            z = Mx[:n, n].clone()
        return R, z, kappa_estimate(At, R, self.lanczos_steps)

    def __call__(self, A, b, delta, tol, iter_lim, rng, logging=True, probs=None):
        if delta != 0:
            raise NotImplementedError("SPOLearned solves the unregularised problem (delta = 0)")
        At = ops.as_device(A).contiguous()
        bt = ops.as_device(b).reshape(-1).contiguous()
        rng = np.random.default_rng(rng)
        dev = At.device
        t0 = time.time()
        R, z, kappa = self.sketch_factor(At, bt, rng, probs)
        if not np.isfinite(kappa) or kappa > self.kappa_max:
            x, log = self.fallback(At, bt, 0.0, tol, iter_lim, rng, logging=logging)
            log.fallback, log.kappa_est = True, kappa
            return ops.like_input(ops.as_device(x), A), log
        log = SketchAndPrecondLog()
        log.time_factor = time.time() - t0
        x0 = ops.trsv_(R, z.clone())
        res = ops.gemv(At, x0, y=bt.clone(), alpha=1.0, beta=-1.0)
        if float((ops.nrm2(res) / ops.nrm2(bt)).item()) >= 1:
            z = None
        tic = time.time()
        out = self.iterative_solver(At, bt, None, 0.0, tol, iter_lim, R, True, z, want_y=False)
        torch.cuda.synchronize(dev)
        log.time_iterate = time.time() - tic
        log.wrap_up(out[2], float("nan"))
        log.fallback, log.kappa_est = False, kappa
        log.error_desc = self.iterative_solver.ERROR_METRIC_INFO
        return ops.like_input(ops.as_device(out[0]), A), log
