"""Surrogate triage of the ensemble: a dense network on a short prefix of the tangent stream predicts
whether an orbit is chaotic (and its log lambda), so that only ambiguous orbits run to full length.

Pipeline (all on the GPU, through the library's kernels; torch only allocates and orders streams):
 1. prefix: every orbit runs `prefix` steps of the standard map with its own K, MEGNO + FLI on one tangent
    frame and the QR spectrum (MGS) on a second; the kernel records per step MEGNO's Y, FLI_B and the log
    stretch (3 x prefix FP32 features, feature-major);
 2. network: features normalised to BF16, a 768-256-256-256-2 ReLU network on the matrix cores (BF16 weights
    and activations, FP32 accumulation) gives a chaotic logit and a predicted log lambda;
 3. triage: orbits with |logit| >= margin take the network's class; the others continue the QR spectrum from
    the prefix state to N/10 and N steps and are labelled by the reference's rule.
The label rule is the reference's P1 instability axis (src/Analysis/labels.jl:6-17): chaotic when
lambda(N) >= 1e-4 and 10 lambda(N) / lambda(N/10) >= 5 (parameters of the study,
workspaces/study/stages/04-analyze/references/p1-le-benchmark-analysis-parameters.toml:51-52); "none detected"
when both fail; "unresolved" otherwise (counted as not chaotic in the binary label).
"""
import ctypes
import math
import time

import numpy as np

from . import _lib as L
from .inputs import torus_uniform, uniform_k
from .reducers import MEGNO, FLI, QRSpectrum, MGS
from .systems import StandardMapK
from .tangent import _context, make_chain

LAMBDA_THRESHOLD = 1e-4
GROWTH_RATIO = 5.0
SIZES = (768, 256, 256, 256, 2)


def to_bf16_bits(a):
    """float32 -> bf16 bit patterns, round to nearest even (the kernels' rounding)."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def p1_labels(lam_N, lam_N10):
    """1 chaotic (positive), 0 none detected, -1 unresolved (src/Analysis/labels.jl:8-15)."""
    lam_N, lam_N10 = np.asarray(lam_N), np.asarray(lam_N10)
    pos10 = lam_N10 > 0
    r = np.where(pos10, 10 * lam_N / np.where(pos10, lam_N10, 1.0), np.nan)
    out = np.full(lam_N.shape, -1, dtype=np.int8)
    out[pos10 & (lam_N >= LAMBDA_THRESHOLD) & (r >= GROWTH_RATIO)] = 1
    out[pos10 & (lam_N < LAMBDA_THRESHOLD) & (r < GROWTH_RATIO)] = 0
    return out


def _check(rc, lib, ctx):
    if rc:
        raise L.HIPError(rc, lib.cl_last_error(ctx).decode())


class SurrogateMLP:
    """768-256-256-256-2 ReLU network: FP32 master weights, BF16 working copies, Adam moments, on the device."""

    def __init__(self, seed=0, sizes=SIZES):
        import torch
        self.lib, self.ctx = _context()
        self.sizes = sizes
        rng = np.random.default_rng(seed)
        dev = torch.device("cuda")
        self.W, self.Wb, self.b, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], [], []
        for i in range(len(sizes) - 1):
            fan_in, fan_out = sizes[i], sizes[i + 1]
            w = rng.standard_normal((fan_out, fan_in)) * math.sqrt(2.0 / fan_in)
            W = torch.from_numpy(w.astype(np.float32)).to(dev)
            self.W.append(W)
            self.Wb.append(torch.from_numpy(to_bf16_bits(w).view(np.int16)).to(dev))
            self.b.append(torch.zeros(fan_out, dtype=torch.float32, device=dev))
            self.mW.append(torch.zeros_like(W))
            self.vW.append(torch.zeros_like(W))
            self.mb.append(torch.zeros(fan_out, dtype=torch.float32, device=dev))
            self.vb.append(torch.zeros(fan_out, dtype=torch.float32, device=dev))
        self.t = 0
        self.mean = None
        self.istd = None

    @property
    def nlayers(self):
        return len(self.sizes) - 1

    def gemm(self, tA, tB, M, N, K, A, lda, B, ldb, C, ldc, f32, bias=None, epi=0, H=None, ldh=0):
        import torch
        st = torch.cuda.current_stream().cuda_stream
        _check(self.lib.cl_mlp_gemm(self.ctx, tA, tB, M, N, K, A, lda, B, ldb, C, ldc, int(f32),
                                    None if bias is None else bias, epi, H, ldh, st), self.lib, self.ctx)

    def forward(self, X, Bn, ldx=None):
        """X: device bf16 (768 x ldx) feature-major, first Bn columns used. Returns (acts, out): acts[l] bf16
        (size_l x Bn), out FP32 (2 x Bn)."""
        import torch
        ldx = Bn if ldx is None else ldx
        acts = []
        A, lda = X.data_ptr(), ldx
        for l in range(self.nlayers):
            M, K = self.sizes[l + 1], self.sizes[l]
            last = l == self.nlayers - 1
            if last:
                out = torch.empty((M, Bn), dtype=torch.float32, device="cuda")
                self.gemm(0, 0, M, Bn, K, self.Wb[l].data_ptr(), K, A, lda, out.data_ptr(), Bn, True,
                          self.b[l].data_ptr(), 2)
                return acts, out
            h = torch.empty((M, Bn), dtype=torch.int16, device="cuda")
            self.gemm(0, 0, M, Bn, K, self.Wb[l].data_ptr(), K, A, lda, h.data_ptr(), Bn, False, self.b[l].data_ptr(), 1)
            acts.append(h)
            A, lda = h.data_ptr(), Bn

    def backward(self, X, Bn, acts, dout, ldx=None):
        """dout bf16 (2 x Bn). Returns FP32 gradients (gW, gb) per layer."""
        import torch
        ldx = Bn if ldx is None else ldx
        gW, gb = [None] * self.nlayers, [None] * self.nlayers
        dZ = dout
        for l in reversed(range(self.nlayers)):
            M, K = self.sizes[l + 1], self.sizes[l]
            Aprev, ldprev = (X.data_ptr(), ldx) if l == 0 else (acts[l - 1].data_ptr(), Bn)
            g = torch.empty((M, K), dtype=torch.float32, device="cuda")
            # dW = dZ (M x B) A_prev^T (B x K): op(B) stored K x B = N x K
            self.gemm(0, 1, M, K, Bn, dZ.data_ptr(), Bn, Aprev, ldprev, g.data_ptr(), K, True)
            gW[l] = g
            db = torch.empty(M, dtype=torch.float32, device="cuda")
            _check(self.lib.cl_mlp_rowsum(self.ctx, M, Bn, dZ.data_ptr(), Bn, db.data_ptr(),
                                          torch.cuda.current_stream().cuda_stream), self.lib, self.ctx)
            gb[l] = db
            if l > 0:
                # dA = W^T (K x M) dZ (M x B), masked by the ReLU of the previous activation
                dA = torch.empty((K, Bn), dtype=torch.int16, device="cuda")
                self.gemm(1, 0, K, Bn, M, self.Wb[l].data_ptr(), K, dZ.data_ptr(), Bn, dA.data_ptr(), Bn, False,
                          None, 3, acts[l - 1].data_ptr(), Bn)
                dZ = dA
        return gW, gb

    def loss(self, out, y, t, w, wreg):
        import torch
        Bn = out.shape[1]
        dout = torch.empty((2, Bn), dtype=torch.int16, device="cuda")
        ng = (Bn + 255) // 256
        part = torch.empty(2 * ng, dtype=torch.float32, device="cuda")
        _check(self.lib.cl_mlp_loss(self.ctx, Bn, out.data_ptr(), y.data_ptr(), t.data_ptr(), w.data_ptr(), wreg,
                                    dout.data_ptr(), part.data_ptr(), torch.cuda.current_stream().cuda_stream),
               self.lib, self.ctx)
        return dout, part

    def adam(self, gW, gb, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        import torch
        self.t += 1
        st = torch.cuda.current_stream().cuda_stream
        for l in range(self.nlayers):
            _check(self.lib.cl_mlp_adam(self.ctx, self.W[l].numel(), self.W[l].data_ptr(), self.Wb[l].data_ptr(),
                                        gW[l].data_ptr(), self.mW[l].data_ptr(), self.vW[l].data_ptr(), self.t, lr, b1,
                                        b2, eps, wd, st), self.lib, self.ctx)
            _check(self.lib.cl_mlp_adam(self.ctx, self.b[l].numel(), self.b[l].data_ptr(), None, gb[l].data_ptr(),
                                        self.mb[l].data_ptr(), self.vb[l].data_ptr(), self.t, lr, b1, b2, eps, 0.0, st),
                   self.lib, self.ctx)

    def _net(self):
        """The cl_mlp_net descriptor of this network's device tensors (their addresses do not change)."""
        if getattr(self, "_netd", None) is None:
            n = L.cl_mlp_net()
            n.nlayers = self.nlayers
            for i, v in enumerate(self.sizes):
                n.sizes[i] = v
            for l in range(self.nlayers):
                for nm, lst in (("W", self.W), ("Wb", self.Wb), ("b", self.b), ("mW", self.mW), ("vW", self.vW),
                                ("mb", self.mb), ("vb", self.vb)):
                    getattr(n, nm)[l] = lst[l].data_ptr()
            self._netd = n
        return self._netd

    def train_step(self, X, ldx, Bn, y, t, w, wreg, lr, loss_sum, b1=0.9, b2=0.999, eps=1e-8):
        """One step on the Bn columns at X (a device address, row stride ldx) with targets at y, t, w (device
        addresses): the library's cl_mlp_train_step, the same kernels as forward, loss, backward and adam called one
        by one, in one call. loss_sum (device FP64, 2) gains the batch-mean loss terms."""
        import torch
        self.t += 1
        _check(self.lib.cl_mlp_train_step(self.ctx, ctypes.byref(self._net()), X, ldx, Bn, y, t, w, wreg, self.t, lr,
                                          b1, b2, eps, loss_sum.data_ptr(), torch.cuda.current_stream().cuda_stream),
               self.lib, self.ctx)

    def set_normalisation(self, trace, Bn):
        import torch
        F = trace.shape[0]
        self.mean = torch.empty(F, dtype=torch.float32, device="cuda")
        self.istd = torch.empty(F, dtype=torch.float32, device="cuda")
        _check(self.lib.cl_mlp_feature_stats(self.ctx, F, Bn, trace.data_ptr(), self.mean.data_ptr(),
                                             self.istd.data_ptr(), torch.cuda.current_stream().cuda_stream),
               self.lib, self.ctx)

    def normalise(self, trace, Bn):
        import torch
        F = trace.shape[0]
        X = torch.empty((F, Bn), dtype=torch.int16, device="cuda")
        _check(self.lib.cl_mlp_normalize(self.ctx, F, Bn, trace.data_ptr(), self.mean.data_ptr(),
                                         self.istd.data_ptr(), X.data_ptr(), torch.cuda.current_stream().cuda_stream),
               self.lib, self.ctx)
        return X

    def train(self, X, y, t, w, epochs=4, batch=4096, lr=1e-3, wreg=0.1, seed=0, log=None, lr_final=0.05,
              native=True):
        """Mini-batch Adam over the columns of X (bf16, 768 x n); y, t, w FP32 device vectors. The step size follows a
        cosine from lr down to lr_final * lr over the run, so the last epochs settle instead of jumping between
        minibatch minima (the network's confidence, and with it the triage's skipped fraction, depends on it).
        `native` runs each step as one cl_mlp_train_step call; False issues its kernels one call at a time (the
        same parameters bit for bit)."""
        import torch
        n = X.shape[1]
        rng = np.random.default_rng(seed)
        hist = []
        total = max(1, epochs * (n // batch))
        it = 0
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).cuda()
            Xp, yp, tp, wp = X[:, perm].contiguous(), y[perm].contiguous(), t[perm].contiguous(), w[perm].contiguous()
            tot = torch.zeros(2, dtype=torch.float64, device="cuda")
            nb = 0
            for s in range(0, n - batch + 1, batch):
                cos = 0.5 * (1.0 + math.cos(math.pi * it / total))
                step_lr = lr * (lr_final + (1.0 - lr_final) * cos)
                if native:  # one library call per step; the batch is a column range of Xp (row stride n)
                    self.train_step(Xp.data_ptr() + 2 * s, n, batch, yp.data_ptr() + 4 * s, tp.data_ptr() + 4 * s,
                                    wp.data_ptr() + 4 * s, wreg, step_lr, tot)
                else:
                    xb = Xp[:, s:s + batch].contiguous()
                    acts, out = self.forward(xb, batch)
                    dout, part = self.loss(out, yp[s:s + batch], tp[s:s + batch], wp[s:s + batch], wreg)
                    gW, gb = self.backward(xb, batch, acts, dout)
                    self.adam(gW, gb, lr=step_lr)
                    tot += part.view(-1, 2).double().sum(0) / batch
                it += 1
                nb += 1
            torch.cuda.synchronize()
            l = (tot / max(nb, 1)).cpu().numpy()
            hist.append(l)
            if log:
                log(f"epoch {ep + 1}/{epochs}: bce {l[0]:.4f} reg {l[1]:.4f}")
        return hist

    def predict(self, X, Bn, chunk=1 << 18):
        import torch
        outs = []
        for s in range(0, Bn, chunk):
            e = min(Bn, s + chunk)
            xb = X[:, s:e].contiguous()
            _, out = self.forward(xb, e - s)
            outs.append(out)
        return torch.cat(outs, dim=1)


# ----------------------------------------------------------------------------------------------- the ensemble side
class Ensemble:
    """Device buffers of an ensemble of the per-orbit-K standard map: x (3 x m), the indicator frame W and the
    QR frame W2 (4 x m each), the state words of the prefix chain (MEGNO, FLI | QR), codes and steps."""

    def __init__(self, x0, seed, elt=1):
        import torch
        self.lib, self.ctx = _context()
        self.elt = elt  # element type of the prefix pass (1 Float64; 2 Float64 orbit, Float32 tangents)
        self.m = m = len(x0)
        self.chain_pre, self.allr = make_chain((MEGNO(replace=False), FLI()), (QRSpectrum(MGS()),))
        self.chain_qr, _ = make_chain((QRSpectrum(MGS()),))
        sd = StandardMapK().descriptor()
        self.sd = sd
        words = self.lib.cl_state_bytes(ctypes.byref(sd), ctypes.byref(self.chain_pre), 1, 2, 2) // 8
        self.w_ind = (MEGNO().nbytes(2, 2) + FLI().nbytes(2, 2)) // 8  # words before the QR state
        dev = torch.device("cuda")
        self.x = torch.from_numpy(np.ascontiguousarray(np.asarray(x0, dtype=np.float64).T)).to(dev)
        self.W = torch.empty((4, m), dtype=torch.float64, device=dev)
        _check(self.lib.cl_initial_deviations_device(self.ctx, seed, 1, m, 2, 2, self.W.data_ptr(),
                                                     torch.cuda.current_stream().cuda_stream), self.lib, self.ctx)
        self.W2 = self.W.clone()
        self.st = torch.zeros((words, m), dtype=torch.int64, device=dev)
        self.code = torch.zeros(m, dtype=torch.int8, device=dev)
        self.h = torch.zeros(m, dtype=torch.int64, device=dev)

    def prefix(self, steps, N):
        """Run the first `steps` steps with the feature trace; returns the FP32 trace (3 steps x m)."""
        import torch
        tr = torch.empty((3 * steps, self.m), dtype=torch.float32, device="cuda")
        rc = self.lib.cl_propagate_device_traced(
            self.ctx, ctypes.byref(self.sd), ctypes.byref(self.chain_pre), self.elt, self.m, 3, 2, 2, 0, steps, N, 1,
            self.x.data_ptr(), self.W.data_ptr(), self.W2.data_ptr(), self.st.data_ptr(), self.code.data_ptr(),
            self.h.data_ptr(), tr.data_ptr(), 0, steps, torch.cuda.current_stream().cuda_stream)
        _check(rc, self.lib, self.ctx)
        return tr


def continue_qr(x, W2, st_qr, code, h, n0, N, steps):
    """Continue the QR spectrum (frame W2 and its state words) of compacted orbits from n0 for `steps` steps."""
    import torch
    lib, ctx = _context()
    m = x.shape[1]
    chain, _ = make_chain((QRSpectrum(MGS()),))
    sd = StandardMapK().descriptor()
    rc = lib.cl_propagate_device(ctx, ctypes.byref(sd), ctypes.byref(chain), 1, m, 3, 2, 2, n0, steps, N, 1,
                                 x.data_ptr(), W2.data_ptr(), None, st_qr.data_ptr(), code.data_ptr(), h.data_ptr(),
                                 torch.cuda.current_stream().cuda_stream)
    _check(rc, lib, ctx)


def qr_lambda1(st_qr):
    """lambda_1 = (sum + comp) / n of QRState words (sums[0].s, sums[0].c, sums[1].s, sums[1].c, n, ...)."""
    import torch
    s = st_qr[0].view(torch.float64)
    c = st_qr[1].view(torch.float64)
    n = st_qr[4].double()
    return torch.where(n > 0, (s + c) / torch.where(n > 0, n, torch.ones_like(n)), torch.zeros_like(n))


def make_starts(seed, m, k_lo=0.5, k_hi=6.0, orbit0=1):
    x = torus_uniform(seed, m, orbit0)
    K = uniform_k(seed, m, k_lo, k_hi, orbit0)
    return np.concatenate([x, K[:, None]], axis=1)


def run_full(x0, seed, N, prefix=256, with_trace=True):
    """The full method on an ensemble: prefix (with features), then the QR spectrum to N/10 and N.
    Returns dict(trace, lam_N10, lam_N, code, times)."""
    import torch
    t0 = time.perf_counter()
    ens = Ensemble(x0, seed)
    tr = ens.prefix(prefix, N)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    w0 = ens.w_ind
    st_qr = ens.st[w0:w0 + 13]
    continue_qr(ens.x, ens.W2, st_qr, ens.code, ens.h, prefix, N, N // 10 - prefix)
    lam10 = qr_lambda1(st_qr).clone()
    continue_qr(ens.x, ens.W2, st_qr, ens.code, ens.h, N // 10, N, N - N // 10)
    lamN = qr_lambda1(st_qr)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return dict(trace=tr if with_trace else None, lam_N10=lam10, lam_N=lamN, code=ens.code, t_prefix=t1 - t0,
                t_rest=t2 - t1)


# ----------------------------------------------------------------------------------------------------- the triage
def targets(run):
    """Binary chaotic label (P1 rule; unresolved counts as not chaotic) and the regression target
    log10(max(lambda(N), 1e-6)), both FP32 device vectors."""
    import torch
    lab = p1_labels(run["lam_N"].cpu().numpy(), run["lam_N10"].cpu().numpy())
    y = torch.from_numpy((lab == 1).astype(np.float32)).cuda()
    t = torch.log10(torch.clamp(run["lam_N"], min=1e-6)).float()
    return y, t, lab


def fit(run, epochs=8, batch=4096, lr=1e-3, wreg=0.1, seed=0, log=None):
    """Train a SurrogateMLP on the prefix features of a full run (the dict of run_full)."""
    import torch
    tr = run["trace"]
    m = tr.shape[1]
    net = SurrogateMLP(seed=seed)
    net.set_normalisation(tr, m)
    X = net.normalise(tr, m)
    y, t, lab = targets(run)
    w = torch.ones(m, dtype=torch.float32, device="cuda")
    hist = net.train(X, y, t, w, epochs=epochs, batch=min(batch, m), lr=lr, wreg=wreg, seed=seed, log=log)
    return net, dict(history=hist, chaotic_fraction=float((lab == 1).mean()))


def calibrate_margin(net, run, target=0.004, grid=(1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 5.0, 6.0, 8.0)):
    """Smallest margin on `grid` whose misclassified fraction on a held-out full run is at most `target`
    (the largest margin of the grid if none is)."""
    tr = run["trace"]
    m = tr.shape[1]
    logit = net.predict(net.normalise(tr, m), m)[0].cpu().numpy()
    _, _, lab = targets(run)
    truth = lab == 1
    for mg in grid:
        conf = np.abs(logit) >= mg
        if np.mean(conf & ((logit > 0) != truth)) <= target:
            return mg
    return grid[-1]


def evaluate(net, run, margin=2.0):
    """Triage outcome on a held-out full run: orbits with |logit| >= margin take the network's class, the rest
    keep the full method's label. Returns the skipped fraction and the misclassified fraction (of all orbits)."""
    tr = run["trace"]
    m = tr.shape[1]
    logit = net.predict(net.normalise(tr, m), m)[0].cpu().numpy()
    _, _, lab = targets(run)
    truth = lab == 1
    conf = np.abs(logit) >= margin
    wrong = conf & ((logit > 0) != truth)
    return dict(skipped_fraction=float(conf.mean()), misclassified=float(wrong.mean()),
                accuracy_all=float(((logit > 0) == truth).mean()), m=int(m))


def triage_run(net, x0, seed, N, prefix=256, margin=2.0, log=None, elt=1):
    """The triaged method on an ensemble: prefix with features, the network, then the QR spectrum to N/10 and
    N for the orbits the network is not sure about. Returns dict(label, lam_N10, lam_N, confident, times);
    lambda entries of confidently triaged orbits are NaN. elt = 2 runs the prefix with Float32 tangents (the
    continuation stays Float64, so the continued orbits' exponents keep the Float64 accuracy)."""
    import torch
    t0 = time.perf_counter()
    m = len(x0)
    ens = Ensemble(x0, seed, elt)
    tr = ens.prefix(prefix, N)
    logit = net.predict(net.normalise(tr, m), m)[0]
    conf = logit.abs() >= margin
    idx = torch.nonzero(~conf).flatten()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mu = int(idx.numel())
    lam10 = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
    lamN = lam10.clone()
    if mu:
        w0 = ens.w_ind
        x = ens.x[:, idx].contiguous()
        W2 = ens.W2[:, idx].contiguous()
        st = ens.st[w0:w0 + 13][:, idx].contiguous()
        code = ens.code[idx].contiguous()
        h = ens.h[idx].contiguous()
        continue_qr(x, W2, st, code, h, prefix, N, N // 10 - prefix)
        lam10[idx] = qr_lambda1(st)
        continue_qr(x, W2, st, code, h, N // 10, N, N - N // 10)
        lamN[idx] = qr_lambda1(st)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    lab = np.where(logit.cpu().numpy() > 0, 1, 0).astype(np.int8)
    conf_h = conf.cpu().numpy()
    if mu:
        full = p1_labels(lamN.cpu().numpy(), lam10.cpu().numpy())
        lab[~conf_h] = (full[~conf_h] == 1)
    if log:
        log(f"triage: {mu} of {m} orbits continue ({mu / m:.3f}); prefix+net {t1 - t0:.2f}s, rest {t2 - t1:.2f}s")
    return dict(label=lab, lam_N10=lam10, lam_N=lamN, confident=conf_h, t_prefix=t1 - t0, t_rest=t2 - t1)


def run_plain(x0, seed, N):
    """The plain method on an ensemble: the QR spectrum (MGS, k = 2) to N/10 and N for every orbit."""
    import torch
    lib, ctx = _context()
    m = len(x0)
    dev = torch.device("cuda")
    x = torch.from_numpy(np.ascontiguousarray(np.asarray(x0, dtype=np.float64).T)).to(dev)
    W = torch.empty((4, m), dtype=torch.float64, device=dev)
    _check(lib.cl_initial_deviations_device(ctx, seed, 1, m, 2, 2, W.data_ptr(),
                                            torch.cuda.current_stream().cuda_stream), lib, ctx)
    st = torch.zeros((13, m), dtype=torch.int64, device=dev)
    code = torch.zeros(m, dtype=torch.int8, device=dev)
    h = torch.zeros(m, dtype=torch.int64, device=dev)
    continue_qr(x, W, st, code, h, 0, N, N // 10)
    lam10 = qr_lambda1(st).clone()
    continue_qr(x, W, st, code, h, N // 10, N, N - N // 10)
    return dict(lam_N10=lam10, lam_N=qr_lambda1(st), code=code)
