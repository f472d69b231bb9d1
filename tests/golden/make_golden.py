"""Write the golden vectors in tests/golden/slab_golden.npz from the reference.

Run in a container that holds the reference checkout (read-only) at
/root/reference:

    python tests/golden/make_golden.py

It covers only what no closed form, invariant or brute-force check pins:

* slab_rpp_periodic and slab_reflected_field (sign and normalisation
  conventions, cubic_scattering/slab_scattering.py:589-735) on seeded fields;
* the S-wave polarisation choice of _build_slab_incident_field
  (slab_scattering.py:450-461) at oblique incidence;
* compute_slab_scattering's converged psi on two tiny seeded slabs
  (aperiodic and periodic), solved to rtol 1e-13;
* compute_cube_tmatrix's scalars (Gamma0, A, B, C, T1-3, amplifications,
  effective contrasts) at the configurations' frequencies: the reference's
  Gamma0 keeps only the static term plus purely imaginary Taylor terms, so no
  exact integral of the Kupradze tensor reproduces its digits.

The fixture stores inputs and outputs only; nothing from the reference is
kept in the repository.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF_ROOT = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "slab_golden.npz"


def main() -> None:
    sys.path.insert(0, str(REF_ROOT))
    from cubic_scattering.effective_contrasts import ReferenceMedium
    from cubic_scattering.slab_scattering import (
        SlabGeometry,
        SlabMaterial,
        SlabResult,
        _build_slab_incident_field,
        compute_slab_scattering,
        compute_slab_tmatrices,
        slab_reflected_field,
        slab_rpp_periodic,
    )

    out: dict[str, np.ndarray] = {}
    ref = ReferenceMedium(alpha=3000.0, beta=1732.0, rho=2500.0)
    out["ref"] = np.array([ref.alpha, ref.beta, ref.rho])

    def material(M, Nz, seed, phi, c):
        rng = np.random.default_rng(seed)
        shape = (Nz, M, M)
        mask = rng.random(shape) < phi
        Dl = np.where(mask, rng.uniform(-c, c, shape) * ref.lam, 0.0)
        Dm = np.where(mask, rng.uniform(-c, c, shape) * ref.mu, 0.0)
        Dr = np.where(mask, rng.uniform(-c, c, shape) * ref.rho, 0.0)
        return Dl, Dm, Dr

    # 1. outputs on seeded fields
    cases = [(4, 2, 0.0, [1.0, 0.0, 0.0]), (5, 3, 1.5e-4, None), (4, 3, 2.5e-4, None)]
    for ci, (M, Nz, p, kh) in enumerate(cases):
        a = 1.0
        omega = 0.4 * ref.beta / a
        if kh is None:
            kh = [np.sqrt(1.0 - (p * ref.alpha) ** 2), p * ref.alpha, 0.0]
        Dl, Dm, Dr = material(M, Nz, 100 + ci, 0.5, 0.3)
        g = SlabGeometry(M=M, N_z=Nz, a=a)
        mat = SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=ref)
        T = compute_slab_tmatrices(g, mat, omega)
        rng = np.random.default_rng(200 + ci)
        psi = rng.standard_normal((Nz, M, M, 9)) + 1j * rng.standard_normal((Nz, M, M, 9))
        res = SlabResult(psi=psi, psi0=psi, geometry=g, material=mat, omega=omega,
                         k_hat=np.asarray(kh, float), wave_type="P", n_gmres_iter=0,
                         gmres_residual=0.0)
        key = f"out{ci}"
        out[f"{key}_dims"] = np.array([M, Nz])
        out[f"{key}_scal"] = np.array([a, omega, p])
        out[f"{key}_khat"] = np.asarray(kh, float)
        out[f"{key}_mat"] = np.stack([Dl, Dm, Dr])
        out[f"{key}_psi"] = psi
        out[f"{key}_rpp"] = np.array([slab_rpp_periodic(res, T, p=p)])
        out[f"{key}_far"] = np.array(slab_reflected_field(res, T))

    # 2. S-wave polarisation at oblique incidence
    g = SlabGeometry(M=2, N_z=2, a=1.0)
    for ci, kh in enumerate([[0.8, 0.36, 0.48], [0.6, 0.0, 0.8], [1.0, 0.0, 0.0]]):
        out[f"swave{ci}_khat"] = np.asarray(kh, float)
        out[f"swave{ci}_psi0"] = _build_slab_incident_field(g, 500.0, ref, np.asarray(kh), "S")

    # 3. converged psi on tiny seeded slabs
    for ci, (M, Nz, periodic) in enumerate([(4, 2, False), (4, 2, True)]):
        a = 1.0
        omega = 0.4 * ref.beta / a
        p = 1.5e-4
        kh = np.array([np.sqrt(1.0 - (p * ref.alpha) ** 2), p * ref.alpha, 0.0])
        Dl, Dm, Dr = material(M, Nz, 300 + ci, 0.3, 0.2)
        g = SlabGeometry(M=M, N_z=Nz, a=a)
        mat = SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=ref)
        res = compute_slab_scattering(g, mat, omega, kh, "P", gmres_tol=1e-13, max_iter=500,
                                      periodic=periodic)
        key = f"solve{ci}"
        out[f"{key}_dims"] = np.array([M, Nz, int(periodic)])
        out[f"{key}_scal"] = np.array([a, omega, p])
        out[f"{key}_khat"] = kh
        out[f"{key}_mat"] = np.stack([Dl, Dm, Dr])
        out[f"{key}_psi"] = res.psi
        out[f"{key}_res"] = np.array([res.gmres_residual])

    # 4. cube T-matrix scalars at the configuration frequencies
    from cubic_scattering.effective_contrasts import MaterialContrast, compute_cube_tmatrix

    fields = ["Gamma0", "Ac", "Bc", "Cc", "T1c", "T2c", "T3c", "amp_u", "amp_theta", "amp_e_off",
              "amp_e_diag", "Drho_star", "Dlambda_star", "Dmu_star_off", "Dmu_star_diag"]
    rows = []
    inputs = []
    for ka, c in [(0.2, 0.1), (0.4, -0.2), (0.75, 0.4), (1.0, 0.3), (0.05, 0.8)]:
        a = 1.0
        omega = ka * ref.beta / a
        con = MaterialContrast(Dlambda=c * ref.lam, Dmu=-0.5 * c * ref.mu, Drho=0.7 * c * ref.rho)
        r = compute_cube_tmatrix(omega, a, ref, con)
        inputs.append([omega, a, con.Dlambda, con.Dmu, con.Drho])
        rows.append([complex(getattr(r, f)) for f in fields])
    out["cube_inputs"] = np.array(inputs)
    out["cube_values"] = np.array(rows)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
