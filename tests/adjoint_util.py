"""Dense FP64 reference for gradients of T: torch autograd through a dense solve.

T = Re Tr[Gamma_L G_LR Gamma_R G_LR^H] with G_LR the contact rows/columns of
A^-1, A = (w + i eta_dev)^2 - D - Sigma_L - Sigma_R assembled densely from the
device blocks (torch leaves) and the oracle's self-energies and contact indices
(negf.py:141-250, :306-357).  complex128 on the CPU.
"""

import numpy as np
import torch

from oracle import negf_oracle as o


def dense_transmission(omega, onsite, coupling, lead_l, lead_r, **kw):
    """T as a torch scalar differentiable w.r.t. the torch tensors ``onsite`` (L, n, n)
    and ``coupling`` (L-1, n, n)."""
    ons_np = [b for b in onsite.detach().numpy()]
    cpl_np = [b for b in coupling.detach().numpy()]
    _, sl, sr, il, ir = o.system(omega, ons_np, cpl_np, lead_l, lead_r, **kw)
    s = kw.get("omega_scale")
    s = 1.0 if s is None else float(s)
    eta = kw.get("eta", 1e-8)
    eta_dev = kw.get("eta_device")
    eta_dev = eta if eta_dev is None else eta_dev
    w, ed = omega / s, eta_dev / s
    ons = onsite.to(torch.complex128) / (s * s)
    cpl = coupling.to(torch.complex128) / (s * s)
    nl, n = ons.shape[0], ons.shape[1]
    dim = nl * n
    d = torch.zeros((dim, dim), dtype=torch.complex128)
    for i in range(nl):
        d[i * n:(i + 1) * n, i * n:(i + 1) * n] = ons[i]
    for i in range(nl - 1):
        d[i * n:(i + 1) * n, (i + 1) * n:(i + 2) * n] = cpl[i]
        d[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n] = cpl[i].conj().T
    z = complex(w, ed) ** 2
    a = z * torch.eye(dim, dtype=torch.complex128) - d
    sl_t = torch.as_tensor(sl)
    sr_t = torch.as_tensor(sr)
    il_t = torch.as_tensor(il)
    ir_t = torch.as_tensor(ir)
    emb_l = torch.zeros((dim, dim), dtype=torch.complex128)
    emb_l[il_t[:, None], il_t[None, :]] = sl_t
    emb_r = torch.zeros((dim, dim), dtype=torch.complex128)
    emb_r[ir_t[:, None], ir_t[None, :]] = sr_t
    a = a - emb_l - emb_r
    rhs = torch.zeros((dim, len(ir)), dtype=torch.complex128)
    rhs[ir_t, torch.arange(len(ir))] = 1.0
    g_cols = torch.linalg.solve(a, rhs)
    glr = g_cols[il_t, :]
    gam_l = 1j * (sl_t - sl_t.conj().T)
    gam_r = 1j * (sr_t - sr_t.conj().T)
    return torch.real(torch.trace(gam_l @ glr @ gam_r @ glr.conj().T))


def dense_grads(omega, onsite, coupling, lead_l, lead_r, **kw):
    """(T, dT/d onsite, dT/d coupling) by autograd, torch's convention d/dRe + i d/dIm."""
    ons = torch.as_tensor(np.asarray(onsite, dtype=np.complex128)).clone().requires_grad_(True)
    cpl = torch.as_tensor(np.asarray(coupling, dtype=np.complex128)).clone().requires_grad_(True)
    t = dense_transmission(omega, ons, cpl, lead_l, lead_r, **kw)
    t.backward()
    g_cpl = cpl.grad.numpy() if cpl.grad is not None else np.zeros(cpl.shape, np.complex128)
    return float(t.item()), ons.grad.numpy(), g_cpl
