"""Child process for the pipeline-variant GPU test (run by
test_gpu_hamiltonian.py with INQAMD_STREAMS / INQAMD_BAND_GROUP_SIZE /
INQAMD_BAND_CHUNK set before the library's first call, since the library
reads them once per process).

On 64^3 x 256 (config 1): Hpsi at ld = 256 (interleaved band groups) against
the same bands at ld = 264 (contiguous chunks), bitwise; an eager ETRS step
likewise; a captured ETRS step replayed twice against the eager contiguous
steps, bitwise; and two bands of Hpsi against the oracle at 1e-12.  Prints
ENV-VARIANT-OK on success.
"""
import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "inq-wannier-fork_amd"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import inq_amd  # noqa: E402
import oracle as O  # noqa: E402
from inq_amd import systems  # noqa: E402


def main():
    env = {k: os.environ.get(k) for k in ("INQAMD_STREAMS", "INQAMD_BAND_GROUP_SIZE",
                                          "INQAMD_BAND_CHUNK")}
    print("[variant]", env, flush=True)
    dev = torch.device("cuda:0")
    s = systems.diamond_system(2)
    nst, ldp = 256, 264
    g = torch.Generator(device=dev).manual_seed(11)
    d = torch.randn(tuple(s.n) + (nst,), dtype=torch.complex128, device=dev, generator=g)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    pad = torch.zeros(tuple(s.n) + (ldp,), dtype=torch.complex128, device=dev)
    pad[..., :nst] = d
    hp = torch.zeros_like(pad)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    h = ham(d)
    ham.handle.call("inqamd_hamiltonian_apply", ptr(pad), ptr(hp), nst, ldp)
    torch.cuda.synchronize()
    assert torch.equal(h, hp[..., :nst]), "grouped Hpsi differs from contiguous"
    assert torch.count_nonzero(hp[..., nst:]).item() == 0
    print("[variant] Hpsi grouped == contiguous", flush=True)
    two = h[..., [0, 200]].cpu().numpy()
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(
        np.ascontiguousarray(d[..., [0, 200]].cpu().numpy()))
    err = float(np.linalg.norm(two - ref) / np.linalg.norm(ref))
    print("[variant] Hpsi vs oracle on 2 bands %.2e" % err, flush=True)
    assert err < 1e-12
    A0 = (ctypes.c_double * 3)(0.01, -0.02, 0.03)
    A1 = (ctypes.c_double * 3)(0.012, -0.018, 0.031)
    for t, ld in ((d, nst), (pad, ldp)):
        ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.02), A0, A1, 4, ptr(t), nst, ld)
    torch.cuda.synchronize()
    assert torch.equal(d, pad[..., :nst]), "grouped ETRS step differs"
    print("[variant] ETRS grouped == contiguous", flush=True)
    zero = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    graph = inq_amd.real_time.etrs_graph(ham, d, 0.02, 4, pert=None)
    for _ in range(2):
        graph.launch(0.0)
        ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.02), zero, zero, 4, ptr(pad), nst,
                        ldp)
        torch.cuda.synchronize()
        assert torch.equal(d, pad[..., :nst]), "captured step differs from eager"
    print("[variant] captured ETRS == eager", flush=True)
    graph.close()
    ham.close()
    print("ENV-VARIANT-OK", flush=True)


if __name__ == "__main__":
    main()
