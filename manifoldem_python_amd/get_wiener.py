"""Drop-in for ``get_wiener`` (``get_wiener.py:9-21``): W_i = sum_{ii<ConOrder}
CTF1[ConOrder-ii+i]^2 + 1/5, a sliding window of squared CTFs, formed on the GPU by a cumulative
sum over the psi-sorted CTF stack."""
import numpy as np
import torch

from . import pipeline as pl


def op(CTF, posPath, posPsi1, ConOrder, num):
    dev = pl.device()
    CTF1 = np.asarray(CTF)[np.asarray(posPath)[np.asarray(posPsi1)], :, :]
    C2 = torch.as_tensor(CTF1, dtype=torch.float64, device=dev) ** 2
    cs = torch.cumsum(C2, dim=0)
    m = num - ConOrder
    # sum_{t=i+1}^{i+ConOrder} C_t^2 = cs[i+ConOrder] - cs[i]
    W = cs[ConOrder:ConOrder + m] - cs[0:m] + 1.0 / 5
    return W.cpu().numpy(), CTF1
