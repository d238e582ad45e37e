"""Child process for the Taylor-chain GPU test (test_gpu_propagate.py runs it
with and without INQAMD_TAYLOR_CHAIN=0, which the library reads once per
process): runs the Taylor sums of the propagators on seeded inputs and saves
the results to the .npz path given as argv[1].

  - 48^3 x 64 bands, 8 atoms, k + A != 0: exponential_in_place (order 4),
    exponential_2_for_1 (both outputs) and two ETRS steps;
  - 64^3 x 256 bands, 64 atoms (4 KB rows, interleaved band groups): one ETRS
    step, and a captured ETRS step replayed twice;
  - the interacting ETRS step (LDA, self-consistent second half) at 48^3.
Prints CHAIN-DUMP-OK on success.
"""
import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "inq-wannier-fork_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import inq_amd  # noqa: E402
from inq_amd import operations, real_time, systems, tddft  # noqa: E402


def main():
    out = {}
    dev = torch.device("cuda:0")
    print("[chain-dump] INQAMD_TAYLOR_CHAIN =", os.environ.get("INQAMD_TAYLOR_CHAIN"), flush=True)
    s = systems.diamond_system(1, grid=48)
    ham = inq_amd.ks_hamiltonian(s, max_bands=64)
    ham.uniform_vector_potential = (0.03, -0.01, 0.02)
    g = torch.Generator(device=dev).manual_seed(5)
    psi = torch.randn(tuple(s.n) + (64,), dtype=torch.complex128, device=dev, generator=g)
    a = psi.clone()
    operations.exponential_in_place(ham, 0.01j, a)
    out["exp48"] = a.cpu().numpy()
    b = psi.clone()
    e = operations.exponential_2_for_1(ham, 0.005j, 0.01j, b)
    out["two48_phi"] = b.cpu().numpy()
    out["two48_exp"] = e.cpu().numpy()
    pert = real_time.fourier_series_vector_potential(np.array([0.02, -0.01, 0.03]), 0.1)
    c = psi.clone()
    for step in range(2):
        real_time.etrs(ham, c, step * 0.01, 0.01, pert)
    out["etrs48"] = c.cpu().numpy()
    print("[chain-dump] 48^3 done", flush=True)
    ham.close()

    s = systems.diamond_system(2)
    ham = inq_amd.ks_hamiltonian(s, max_bands=256)
    g = torch.Generator(device=dev).manual_seed(6)
    psi = torch.randn(tuple(s.n) + (256,), dtype=torch.complex128, device=dev, generator=g)
    psi *= 1.0 / np.sqrt(s.volume)
    c = psi.clone()
    real_time.etrs(ham, c, 0.0, 0.02, pert)
    out["etrs64"] = c.cpu().numpy()
    d = psi.clone()
    graph = real_time.etrs_graph(ham, d, 0.02, 4, pert=None)
    for _ in range(2):
        graph.launch(0.0)
    torch.cuda.synchronize()
    graph.close()
    out["graph64"] = d.cpu().numpy()
    print("[chain-dump] 64^3 done", flush=True)
    ham.close()

    s = systems.diamond_system(1, grid=48)
    ham = inq_amd.ks_hamiltonian(s, max_bands=8)
    psi = torch.from_numpy(systems.random_orbitals(s, 8)).to(dev)
    vion = torch.from_numpy(np.ravel(s.vloc).copy()).to(dev)
    tddft.ks_potential(ham, vion, tddft.density(ham, psi, 2.0), xc="lda")
    nscf = [tddft.etrs(ham, psi, vion, step * 0.02, 0.02, 2.0, pert, xc="lda") for step in range(2)]
    out["tddft48"] = psi.cpu().numpy()
    out["tddft48_nscf"] = np.array(nscf)
    print("[chain-dump] interacting done", flush=True)
    ham.close()
    np.savez(sys.argv[1], **out)
    print("CHAIN-DUMP-OK", flush=True)


if __name__ == "__main__":
    main()
