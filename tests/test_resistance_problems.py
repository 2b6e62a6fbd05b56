"""The strain problems of the resistance driver (StokesInlet3sidesTorquePiCRS, src/lib_particle.f90:6-168)
and closed forms of the ellipsoid that pin the oracle's rigid-body and strain problems.

Closed forms for a triaxial ellipsoid with semi-axes a_i (Kim & Karrila, Microhydrodynamics, ch. 3;
Oberbeck 1876, Jeffery 1922), with chi = int_0^inf dl / D(l), alpha_i = int_0^inf dl / ((a_i^2 + l) D(l)),
D(l) = sqrt(prod_j (a_j^2 + l)), evaluated with Carlson's integrals (chi = 2 R_F(a1^2, a2^2, a3^2),
alpha_i = (2/3) R_D(a_j^2, a_k^2, a_i^2)), viscosity 1:
  translation   F_i = 16 pi / (chi + a_i^2 alpha_i)
  rotation      T_i = (16 pi / 3) (a_j^2 + a_k^2) / (a_j^2 alpha_j + a_k^2 alpha_k)
  strain        T_i = (16 pi / 3) (a_j^2 - a_k^2) / (a_j^2 alpha_j + a_k^2 alpha_k) for the unit strain
                E_jk = E_kj = 1 ((0, z, y) for i = x, (z, 0, x) for i = y, (y, x, 0) for i = z);
the strain fields exert no force. For a sphere they reduce to 6 pi a, 8 pi a^3 and 0.
"""
import numpy as np
import pytest
from scipy.special import elliprd, elliprf

from andromeda_amd import api
from andromeda_amd import mesh as M
from andromeda_amd.metrics import tensor_error

AXES = np.array([1.0, 0.7, 0.5])


def ellipsoid_closed_forms(a):
    a2 = np.asarray(a, dtype=np.float64) ** 2
    chi = 2.0 * elliprf(a2[0], a2[1], a2[2])
    al = np.array([2.0 / 3.0 * elliprd(a2[(i + 1) % 3], a2[(i + 2) % 3], a2[i]) for i in range(3)])
    j, k = [(i + 1) % 3 for i in range(3)], [(i + 2) % 3 for i in range(3)]
    den = a2[j] * al[j] + a2[k] * al[k]
    K = 16 * np.pi / (chi + a2 * al)
    Om = 16 * np.pi / 3 * (a2[j] + a2[k]) / den
    Pi = 16 * np.pi / 3 * (a2[j] - a2[k]) / den
    return K, Om, Pi


def test_closed_forms_reduce_to_sphere():
    K, Om, Pi = ellipsoid_closed_forms([0.8, 0.8, 0.8])
    assert np.allclose(K, 6 * np.pi * 0.8) and np.allclose(Om, 8 * np.pi * 0.8 ** 3) and np.allclose(Pi, 0.0)


def _oracle_errors(orc, nu):
    m = M.superellipsoid(AXES, 2.0, nu)
    K, Om, Pi = ellipsoid_closed_forms(AXES)
    FT, _, _ = orc.resistance(m, exterior=True, eps=1e-12)
    FS, _, _ = orc.resistance(m, exterior=True, eps=1e-12, problems=1)
    eK = np.abs(np.diag(FT[:3, :3]) / K - 1).max()
    eO = np.abs(np.diag(FT[3:, 3:]) / Om - 1).max()
    eP = np.abs(np.array([FS[0, 3], FS[1, 4], FS[2, 5]]) / Pi - 1).max()
    return eK, eO, eP, FS


def test_oracle_ellipsoid_rigid_and_strain_closed_forms(orc):
    """The oracle's translation, rotation and strain problems on a 1 : 0.7 : 0.5 ellipsoid converge
    to the closed forms at second order in the element size (F = 320 -> 720: errors fall by about
    (4/6)^2 = 0.44), and the strain fields give no force and only the diagonal torques."""
    e4 = _oracle_errors(orc, 4)
    e6 = _oracle_errors(orc, 6)
    for a, b in zip(e4[:3], e6[:3]):
        assert b < 0.025 and b < 0.55 * a, (a, b)
    FS = e6[3]
    scale = np.abs(FS[:3, 3:]).max()
    assert np.abs(FS[:3, :3]).max() < 1e-9 * scale
    assert np.abs(FS[:3, 3:] - np.diag(np.diag(FS[:3, 3:]))).max() < 1e-9 * scale
    assert np.all(FS[3:] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [api.FORM_EXTERIOR, api.FORM_BOUNDED])
def test_gpu_strain_plain_vs_oracle(ctx, orc, form):
    if form == api.FORM_EXTERIOR:
        m = M.superellipsoid(AXES, 2.0, 6)
    else:
        m = M.body_in_sphere(M.superellipsoid(AXES, 2.0, 4), 16.0, 2)
    FS, _, _ = orc.resistance(m, exterior=form == api.FORM_EXTERIOR, eps=1e-12, problems=1)
    r = ctx.resistance(m, form=form, precision=api.PREC_FP64, eps=1e-12, problems=api.PROBLEMS_STRAIN)
    assert r.status == 0 and np.all(r.nit[3:] == 0)
    rel = np.abs(r.FT - FS).max() / np.abs(FS).max()
    print(f"form {form}: strain FT vs oracle {rel:.2e}")
    assert rel < 1e-9


@pytest.mark.gpu
def test_gpu_strain_fast_forms_and_closed_form(ctx):
    """Fast forms on the strain problems within 1e-6 of the plain path (torques relative to the
    largest); the plain path's Pi torques within 1 % of Jeffery's at 2000 triangles; a sphere
    gets no strain torque."""
    m = M.superellipsoid(AXES, 2.0, 10)
    p = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12, problems=api.PROBLEMS_STRAIN)
    _, _, Pi = ellipsoid_closed_forms(AXES)
    Tp = np.array([p.FT[0, 3], p.FT[1, 4], p.FT[2, 5]])
    assert np.abs(Tp / Pi - 1).max() < 0.01
    for prec in (api.PREC_MIXED, api.PREC_LEARNED):
        if prec == api.PREC_LEARNED:
            import os
            z = np.load(os.path.join(os.path.dirname(api.__file__), "quad_model.npz"))
            ctx.set_quadrature_model(z["params"], float(z["confidence"]), float(z["near_ratio"]), z["far_table"])
        f = ctx.resistance(m, precision=prec, problems=api.PROBLEMS_STRAIN)
        assert f.status == 0
        err = np.abs(f.FT - p.FT).max() / np.abs(p.FT).max()
        print(f"precision {prec}: strain err {err:.2e}, its {f.nit.tolist()}")
        assert err < 1e-6
    s = ctx.resistance(M.geodesic_sphere(6), precision=api.PREC_FP64, eps=1e-12, problems=api.PROBLEMS_STRAIN)
    r = ctx.resistance(M.geodesic_sphere(6), precision=api.PREC_FP64, eps=1e-12)
    assert np.abs(s.FT).max() < 1e-6 * r.FT[3, 3]
    # the rigid problems of the same body are unchanged by the option
    q = ctx.resistance(m, precision=api.PREC_FP64, eps=1e-12)
    K, Om, _ = ellipsoid_closed_forms(AXES)
    assert np.abs(np.diag(q.FT[:3, :3]) / K - 1).max() < 0.01 and np.abs(np.diag(q.FT[3:, 3:]) / Om - 1).max() < 0.01
    assert tensor_error(q.FT, q.FT.T) < 1e-9
