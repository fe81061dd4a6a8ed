"""Radiation-coupling cases in the reference's unit system (lib/global_consts.h,
CGS branch): the physical constants, the cooling-function table fixture and a
seeded blast of physical states (rho ~ 1e-8 g/cm^3, T ~ 1e4..1e5 K) with
radiation fields whose energy is a few per cent of the gas pressure."""
import lzma
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# lib/global_consts.h:12-83, CGS values
kC_Light = 3e10
kDist = 1e13
kGravity = 6.6726e-8
kM_hydrogen = 1.6735575e-24
k_boltzmann = 1.3807e-16
kStefanBoltzmann = 5.670374419e-5
kSigma_thomson = 6.65210e-25
kMass = kC_Light * kC_Light * kDist / kGravity
kTime = kDist / kC_Light
kDensity = kMass / kDist**3
kPressure = kMass / (kDist * kTime * kTime)
kRadiation = kMass / kTime**3
a_rad = 4.0 * kStefanBoltzmann / kC_Light  # radiation constant, erg cm^-3 K^-4


def table_path(tmp_dir) -> str:
    """tab_func/rcf_H.dat of the reference, decompressed from the fixture into tmp_dir."""
    path = os.path.join(str(tmp_dir), "rcf_H.dat")
    if not os.path.exists(path):
        with lzma.open(os.path.join(GOLDEN, "rcf_H.dat.xz")) as f, open(path, "wb") as g:
            g.write(f.read())
    return path


def code_state(rho_cgs, T, v):
    """W in code units of gas at density rho_cgs (g/cm^3) and temperature T (K), the
    ideal-gas law p = rho k_B T / m_H."""
    rho_cgs, T = np.asarray(rho_cgs, float), np.asarray(T, float)
    d = rho_cgs / kDensity
    p = rho_cgs * k_boltzmann * T / kM_hydrogen / kPressure
    return np.column_stack([d, np.broadcast_to(v, (d.size, 3)), p])


def physical_blast(mesh, seed=3, ur_scale=1.0):
    """W0 (hot sphere of radius 0.25 at 1e5 K in gas at 1e4 K, rho 1e-8..1.3e-8 g/cm^3,
    |v| ~ 1e-6 c) and FP32 radiation moments (Ur = ur_scale * a T_r^4 with T_r in
    [1e4, 2e4] K, |F| ~ 0.3 Ur) per cell."""
    rng = np.random.default_rng(seed)
    n = mesh.n_cells
    r = np.linalg.norm(mesh.cells["center"] - 0.5, axis=1)
    rho = 1e-8 * (1 + 0.3 * rng.random(n))
    T = np.where(r < 0.25, 1e5, 1e4) * (1 + 0.1 * rng.random(n))
    W0 = code_state(rho, T, 0.0)
    W0[:, 1:4] = 1e-6 * rng.normal(size=(n, 3))
    Tr = 1e4 * (1 + rng.random(n))
    Ur = (ur_scale * a_rad * Tr**4 / kPressure).astype(np.float32)
    Fr = (0.3 * Ur[:, None] * rng.normal(size=(n, 3)) / np.sqrt(3)).astype(np.float32)
    return W0, Ur, Fr
