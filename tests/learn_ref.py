"""fp32/fp64 torch reference of the trainable layered min-sum decoder.

Written from the model's definition (independently of the HIP kernels):
per iteration t and schedule step, every check (i, j) of the step's layers
reads x_e = L[v_e] - m_e at the step-start APP, then
    y_e = clamp(alpha[t, i] * s_e * relu(mu_e - beta[t, i]), -127, 127),
mu_e = min of |x| over the check's other edges, s_e = product of the other
edges' signs (x < 0 negative), and L[v_e] += y_e - m_e, m_e = y_e, the
updates of a step summed.  Geometry: address a of table row r is circulant
(r, s = a div q) of layer a mod q; check j of layer i reads info VN
360 r + (j - s) mod 360, parity VN K + q j + i and K + q j + i - 1 (i > 0) or
K + q (j - 1) + q - 1 (i = 0, j > 0); messages are ordered layer by layer,
edge-major (lib/cusat_dsp/cuda/ldpc_decode_kernel.cu:1006-1065).
Loss = mean over iterations of mean(softplus(-kappa * s * L_t)), s = +1 for
a 0 bit.  torch autograd supplies the gradients.
"""
import numpy as np
import torch

from tests.test_tables import SCHEDS, TABLES


class TorchLayered:
    def __init__(self, name, use_sched=True):
        t = TABLES[name]
        N, K = t["N"], t["K"]
        q = (N - K) // 360
        self.N, self.K, self.q = N, K, q
        lay = [[] for _ in range(q)]
        for r, row in enumerate(t["rows"]):
            for a in row:
                lay[a % q].append((r, a // q))
        sc = SCHEDS.get(name) if use_sched else None
        if sc is not None and sc["q"] == q:
            self.steps = [tuple(p) for p in sc["pairs"]]
        else:
            self.steps = [(i, -1) for i in range(q)]
        j = np.arange(360)
        self.V, self.MI, self.MASK = [], [], []
        moff = 0
        for i in range(q):
            cols = [360 * r + (j - s) % 360 for (r, s) in lay[i]]
            cols.append(K + q * j + i)
            if i > 0:
                cols.append(K + q * j + i - 1)
            else:
                cols.append(np.where(j > 0, K + q * (j - 1) + q - 1, 0))
            deg = len(cols)
            V = np.stack(cols, 1)
            mask = np.ones((360, deg), bool)
            if i == 0:
                mask[0, deg - 1] = False
            MI = moff + np.arange(deg)[None, :] * 360 + j[:, None]
            moff += 360 * deg
            self.V.append(V)
            self.MI.append(MI)
            self.MASK.append(mask)
        self.E = moff
        self.D = max(v.shape[1] for v in self.V)

    def _step_tensors(self, layers, device):
        D = self.D
        V, MI, MASK, LAY = [], [], [], []
        for i in layers:
            d = self.V[i].shape[1]
            pad = D - d
            V.append(np.pad(self.V[i], ((0, 0), (0, pad))))
            MI.append(np.pad(self.MI[i], ((0, 0), (0, pad))))
            MASK.append(np.pad(self.MASK[i], ((0, 0), (0, pad))))
            LAY.append(np.full(360, i))
        t = lambda a, dt: torch.as_tensor(np.concatenate(a), dtype=dt, device=device)
        return t(V, torch.long), t(MI, torch.long), t(MASK, torch.bool), t(LAY, torch.long)

    def run(self, llr, bits, alpha, beta, kappa, quant=False):
        """llr [B, N], bits [B, N] (0/1), alpha/beta [T, q] (tensors, may
        require grad).  Returns (final APP, per-iteration losses, loss).

        quant: the integer decoder's check node -- beta_q = max(round(beta), 0),
        n = clamp(round(64 alpha), 1, 1023), |y| = min(floor((r n + 32) / 64),
        127), r = relu(mu - beta_q) -- with straight-through gradients
        dy/dalpha = s r, dy/dbeta = -alpha s [mu > beta_q], dy/dmu = alpha s
        [mu > beta_q], all gated by floor(...) <= 127."""
        B = llr.shape[0]
        T = alpha.shape[0]
        dev, dt = llr.device, llr.dtype
        L = llr.clone()
        M = torch.zeros(B, self.E, dtype=dt, device=dev)
        s = 1.0 - 2.0 * bits.to(dt)
        cache = [self._step_tensors([i for i in st if i >= 0], dev) for st in self.steps]
        inf = torch.tensor(float("inf"), dtype=dt, device=dev)
        ar = torch.arange(self.D, device=dev)
        losses = []
        for t in range(T):
            for V, MI, MASK, LAY in cache:
                mold = M[:, MI]
                x = torch.where(MASK, L[:, V] - mold, inf)
                a = x.abs()
                mn1, i1 = a.min(2)
                a2 = a.scatter(2, i1.unsqueeze(2), float("inf"))
                mn2, _ = a2.min(2)
                mu = torch.where(ar == i1.unsqueeze(2), mn2.unsqueeze(2), mn1.unsqueeze(2))
                neg = (x < 0) & MASK
                par = neg.sum(2, keepdim=True) % 2
                sg = torch.where(((par + neg.long()) % 2) == 1, -1.0, 1.0).to(dt)
                al = alpha[t][LAY].view(1, -1, 1)
                be = beta[t][LAY].view(1, -1, 1)
                if quant:
                    bq = torch.clamp(torch.round(be.detach()), min=0.0)
                    n = torch.clamp(torch.round(64.0 * al.detach()), 1.0, 1023.0)
                    r = torch.relu(mu - bq)
                    m = torch.floor((r * n + 32.0) / 64.0)
                    pas = (m <= 127.0).to(dt)
                    act = (mu > bq).to(dt)
                    yh = sg * torch.clamp(m, max=127.0)
                    f = pas * sg * (al * r.detach() + al.detach() * act * (mu - be))
                    y = yh.detach() + (f - f.detach())
                else:
                    y = torch.clamp(al * sg * torch.relu(mu - be), -127.0, 127.0)
                y = torch.where(MASK, y, torch.zeros((), dtype=dt, device=dev))
                fm = MASK.reshape(-1)
                delta = (y - torch.where(MASK, mold, torch.zeros((), dtype=dt, device=dev))).reshape(B, -1)
                L = L.index_add(1, V.reshape(-1)[fm], delta[:, fm])
                M = M.index_put((torch.arange(B, device=dev)[:, None], MI.reshape(-1)[fm][None, :]),
                                y.reshape(B, -1)[:, fm])
            losses.append(torch.nn.functional.softplus(-kappa * s * L).mean())
        loss = torch.stack(losses).mean()
        return L, torch.stack(losses), loss
