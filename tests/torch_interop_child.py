"""Child process of test_torch_tensors_interoperate (tests/test_gpu_leaf.py).

It imports torch and touches the GPU through it before the library starts,
so torch's HIP runtime is the first in the process, then passes torch ROCm
tensors straight to the leaf gather (picsarpy accepts them in place) and
compares the result with the oracle.  Prints "OK <relative error>"."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import torch  # noqa: E402

NX, NY, NZ, NG = 20, 18, 14, 3
SHAPE = (NX + 1 + 2 * NG, NY + 1 + 2 * NG, NZ + 1 + 2 * NG)


def main():
    if not torch.cuda.is_available():
        print("NOGPU torch sees no GPU")
        return 3
    torch.ones(1, device="cuda").sum().item()  # torch's runtime first
    from oracle import load_oracle
    from picsar_amd import _lib
    from picsar_amd.picsarpy import picsar
    _lib.call("pxrhip_init", 0)
    orc = load_oracle()
    rng = np.random.default_rng(12)
    n = 5000
    x = [rng.uniform(1, 10, n) for _ in range(3)]
    F = [np.asfortranarray(rng.standard_normal(SHAPE)) for _ in range(6)]
    e_ref = [np.zeros(n) for _ in range(6)]
    orc.geteb3d_energy_conserving(*x, *e_ref, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, 2, 2, 2,
                                  *F, True, False)
    tx = [torch.tensor(a, device="cuda") for a in x]
    te = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(6)]
    # grid tensors in both accepted layouts: a Fortran-strided view and a
    # C-contiguous array of the reversed shape
    tF = [torch.tensor(np.ascontiguousarray(f.T), device="cuda").permute(2, 1, 0) for f in F[:3]]
    tF += [torch.tensor(np.ascontiguousarray(f.T), device="cuda") for f in F[3:]]
    picsar.l_nodalgrid = False
    picsar.geteb3d_energy_conserving(n, *tx, *te, 0, 0, 0, 1, 1, 1, NX, NY, NZ, NG, NG, NG, 2, 2,
                                     2, *tF, False, True, 8, 0)
    worst = 0.0
    for a, b in zip(te, e_ref):
        a = a.cpu().numpy()
        worst = max(worst, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    print(f"OK {worst:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
