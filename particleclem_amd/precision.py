"""Error norms between two device levels of the same box table (the
mixed-precision step against the FP64 step), computed on the GPU.

For a quantity a (test) against b (reference) over the valid cells of every
box:  L1 = sum |a - b| / sum |b|,  Linf = max |a - b| / max |b|.
Quantities: rho (U[URHO]), T (U[UTEMP]) and each mass fraction
Y_k = U[UFS + k] / U[URHO].
"""
URHO, UTEMP, UFS = 0, 6, 7


def error_norms(test, ref, nspec, nboxes):
    import torch
    names = ["rho", "T"] + [f"Y{k}" for k in range(nspec)]
    nq = len(names)
    dsum = torch.zeros(nq, dtype=torch.float64, device="cuda")
    rsum = torch.zeros_like(dsum)
    dmax = torch.zeros_like(dsum)
    rmax = torch.zeros_like(dsum)
    for b in range(nboxes):
        a = test.valid_tensor(b)
        r = ref.valid_tensor(b)

        def qty(U):
            out = [U[URHO], U[UTEMP]]
            out += [U[UFS + k] / U[URHO] for k in range(nspec)]
            return torch.stack([x.reshape(-1) for x in out])

        qa, qr = qty(a), qty(r)
        d = (qa - qr).abs()
        dsum += d.sum(dim=1)
        rsum += qr.abs().sum(dim=1)
        dmax = torch.maximum(dmax, d.max(dim=1).values)
        rmax = torch.maximum(rmax, qr.abs().max(dim=1).values)
    l1 = (dsum / rsum).tolist()
    linf = (dmax / rmax).tolist()
    return {n: {"L1": l1[i], "Linf": linf[i]} for i, n in enumerate(names)}
