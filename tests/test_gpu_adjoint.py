"""The adjoint Jacobian leg: level Jacobians from one tangent batch per block
(UpwellingJacobian(method="adjoint"): clblm_od_tangent + clblm_rt_jac +
clblm_level_jac), and the differentiable layer OD (layer_od, a torch.autograd.Function).

* Central differences of the whole OD + radiance pipeline (the same model: closed-form
  line shapes) with the north star's steps, dT = 0.01 K and dln q = 1e-4, perturbing a
  level as drv_LevelTemp / drv_LevelConc linearise it (clblm_noScattJac.f90:1005-1034,
  1159-1171): within 1e-4 of the largest value.  The model keeps the reference's
  discontinuities -- a line's walk ends at 4 half-widths where SHAPEG and VERFN are not
  zero, CNVF4Q switches form at 64 Voigt widths, IZETA steps -- and a step that
  crosses one moves an isolated point by the jump over 2 dT.  Such points are allowed
  (at most 1 in 10^4 of the points) only if they behave as jumps when the step is cut
  4x: their difference error grows (the jump is still crossed) or vanishes (it no
  longer is), where a smooth point's error falls 16x and a wrong derivative's stays.
* The reference's own Jacobians (a one-sided 1 K OD difference for dT, the OD of the
  molecule alone for dln q; UpwellingJacobian(method="variants") in FP64): within the
  O(dT) curvature of the one-sided difference and the self-broadening the
  molecule-alone OD leaves out.
* layer_od's backward pass against central differences of a weighted sum of its output.
"""
import copy

import numpy as np
import pytest
import torch

from clblm_amd import workloads as W
from clblm_amd.jacobian import UpwellingJacobian, layer_od
from clblm_amd.odlay import Context

pytestmark = pytest.mark.gpu

V1, V2 = 640.0, 646.0


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _scene(nlay=8, n_lines=3000):
    lines = W.profile_lines(n_lines, 500.0, 800.0, seed=1001)
    lines.s[:] = lines.s * np.float32(1e-3)
    noise = np.random.default_rng(5).normal(0.0, 3.0, nlay)
    layers = W.atmosphere(nlay, 1.0, t_noise=noise)
    t_lev = np.array([layers[0].tave + 2.0] + [0.5 * (a.tave + b.tave) for a, b in zip(layers[:-1], layers[1:])] +
                     [layers[-1].tave - 1.0])
    p_lev = W.level_pressures(nlay, 1.0)
    q_lev = W.level_densities(nlay, 1.0)
    dvs = [W.calculate_dv(650.0, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    return lines, layers, t_lev, p_lev, q_lev, dvs


def _cont(lay, a, b):
    return W.continuum_absrb(lay, a, b)


def _jac(ctx, handle, scene, mols, method="adjoint", precision=None):
    from clblm_amd._lib import FP64
    _, layers, t_lev, p_lev, q_lev, dvs = scene
    return UpwellingJacobian(ctx, handle, layers, t_lev, dvs, p_lev=p_lev, q_lev=q_lev, mols=mols,
                             precision=FP64 if precision is None else precision, continuum=_cont, method=method)


def _perturbed(ctx, handle, scene, base, k, dt_lev=0.0, dlnq=0.0, mol=None):
    """radiance of the scene with level k's temperature moved by dt_lev (its layers'
    temperatures by dT/dT_lev, their broadening columns by -Wtot dT/T) or molecule
    mol's level density scaled by exp(dlnq) (its layers' columns by the level2layer
    exponents)"""
    lines, layers, t_lev, p_lev, q_lev, dvs = scene
    layers = [copy.deepcopy(l) for l in layers]
    t_lev = t_lev.copy()
    mi = base.mols.index(mol) if mol else 0
    for l in (k - 1, k):
        if l < 0 or l >= len(layers):
            continue
        upper = l == k - 1
        if dt_lev:
            d = (base.dt_dtup[l] if upper else base.dt_dtlo[l]) * dt_lev
            layers[l].wtot = layers[l].wtot * (1.0 - d / layers[l].tave)
            layers[l].tave = layers[l].tave + d
        if dlnq:
            e = base.dw_dqup[mi, l] if upper else base.dw_dqlo[mi, l]
            layers[l].w[mol] = layers[l].w[mol] * np.exp(dlnq * e)
    t_lev[k] += dt_lev
    sc = (lines, layers, t_lev, p_lev, q_lev, dvs)
    return _jac(ctx, handle, sc, base.mols).block(V1, V2)["rad"].cpu().numpy()


@pytest.fixture(scope="module")
def scene_h(ctx):
    sc = _scene(8)
    h = ctx.upload_lines(sc[0])
    yield sc, h
    ctx.free_lines(h)


def _check_fd(jac, fd, fd4, what):
    """jac against the central difference fd (step h) and fd4 (step h/4)"""
    scale = np.abs(fd).max()
    e, e4 = np.abs(jac - fd) / scale, np.abs(jac - fd4) / scale
    bad = e > 1e-4
    print("%s: %d points, median |adjoint - central FD| / max %.1e, above 1e-4: %d (max %.2e; at h/4 %s)" %
          (what, fd.size, np.median(e), int(bad.sum()), e.max(), ["%.1e" % x for x in e4[bad]]))
    assert bad.sum() <= max(3, fd.size // 10000), what
    jump = (e4[bad] > 2.0 * e[bad]) | (e4[bad] < 1e-3 * e[bad])
    assert np.all(jump), what


@pytest.mark.parametrize("k", [0, 3, 8])
def test_dT_matches_central_difference_of_the_pipeline(ctx, scene_h, k):
    sc, h = scene_h
    jb = _jac(ctx, h, sc, [1])
    jac = jb.block(V1, V2)["jac_t"][k].cpu().numpy()
    fds = []
    for d in (0.01, 0.0025):
        fds.append((_perturbed(ctx, h, sc, jb, k, dt_lev=d) - _perturbed(ctx, h, sc, jb, k, dt_lev=-d)) / (2 * d))
    _check_fd(jac, fds[0], fds[1], "dR/dT level %d" % k)


@pytest.mark.parametrize("mol,k", [(2, 2), (3, 6), (1, 4)])
def test_dlnq_matches_central_difference_of_the_pipeline(ctx, scene_h, mol, k):
    sc, h = scene_h
    jb = _jac(ctx, h, sc, [mol])
    jac = jb.block(V1, V2)["jac_lnq"][0][k].cpu().numpy()
    fds = []
    for e in (1e-4, 2.5e-5):
        fds.append((_perturbed(ctx, h, sc, jb, k, dlnq=e, mol=mol) -
                    _perturbed(ctx, h, sc, jb, k, dlnq=-e, mol=mol)) / (2 * e))
    _check_fd(jac, fds[0], fds[1], "dR/dln q molecule %d level %d" % (mol, k))


def test_against_the_references_own_jacobians(ctx, scene_h):
    """The reference's Jacobians (FP64, drv_LevelTemp's T + 1 K one-sided difference,
    drv_LevelConc's molecule-alone OD) against the adjoint's, level by level."""
    sc, h = scene_h
    a = _jac(ctx, h, sc, [1, 2, 3]).block(V1, V2)
    r = _jac(ctx, h, sc, [1, 2, 3], method="variants").block(V1, V2)
    rad_a, rad_r = a["rad"].cpu().numpy(), r["rad"].cpu().numpy()
    # the two ODs differ by the shape tables' spacing only
    drad = np.abs(rad_a - rad_r).max() / np.abs(rad_r).max()
    print("radiance: max |adjoint model - table model| / max = %.2e" % drad)
    assert drad < 1e-3
    jt_a, jt_r = a["jac_t"].cpu().numpy(), r["jac_t"].cpu().numpy()
    err_t = np.abs(jt_a - jt_r).max() / np.abs(jt_r).max()
    jq_a, jq_r = a["jac_lnq"].cpu().numpy(), r["jac_lnq"].cpu().numpy()
    err_q = [np.abs(jq_a[m] - jq_r[m]).max() / np.abs(jq_r[m]).max() for m in range(3)]
    print("dR/dT_lev: max |adjoint - reference| / max = %.2e; dR/dln q: %s" % (err_t, ["%.2e" % e for e in err_q]))
    assert err_t < 0.1
    assert max(err_q) < 0.05


def test_layer_od_backward_matches_central_differences(ctx, scene_h):
    sc, h = scene_h
    lines, layers, t_lev, p_lev, q_lev, dvs = sc
    mols = (1, 2)
    dvout = min(dvs)
    L = len(layers)
    t = torch.tensor([l.tave for l in layers], dtype=torch.float64, requires_grad=True)
    lwt = torch.tensor([np.log(l.wtot) for l in layers], dtype=torch.float64, requires_grad=True)
    lw = torch.tensor([[np.log(l.w[m]) for m in mols] for l in layers], dtype=torch.float64, requires_grad=True)
    ctx.jrad = 1
    od = layer_od(ctx, h, layers, V1, V2, dvs, t, lwt, lw, mols, dvout, continuum=_cont)
    wts = torch.as_tensor(np.random.default_rng(3).uniform(0.5, 1.5, tuple(od.shape)), device=od.device)
    loss = (od * wts).sum()
    loss.backward()

    def f(tt, a, b):
        ctx.jrad = 1
        with torch.no_grad():
            return float((layer_od(ctx, h, layers, V1, V2, dvs, tt, a, b, mols, dvout, continuum=_cont) * wts).sum())

    for l in (0, 4, L - 1):
        for name, x, g, step in (("T", t, t.grad, 0.01), ("lnWtot", lwt, lwt.grad, 1e-4)):
            xp, xm = x.detach().clone(), x.detach().clone()
            xp[l] += step
            xm[l] -= step
            args_p = (xp, lwt.detach(), lw.detach()) if name == "T" else (t.detach(), xp, lw.detach())
            args_m = (xm, lwt.detach(), lw.detach()) if name == "T" else (t.detach(), xm, lw.detach())
            fd = (f(*args_p) - f(*args_m)) / (2 * step)
            err = abs(float(g[l]) - fd) / max(abs(fd), 1e-300)
            print("layer %d d loss / d %s: backward %.6e, central FD %.6e, rel err %.2e" % (l, name, float(g[l]), fd, err))
            assert err < 1e-5
        for k in range(len(mols)):
            xp, xm = lw.detach().clone(), lw.detach().clone()
            xp[l, k] += 1e-4
            xm[l, k] -= 1e-4
            fd = (f(t.detach(), lwt.detach(), xp) - f(t.detach(), lwt.detach(), xm)) / 2e-4
            err = abs(float(lw.grad[l, k]) - fd) / abs(fd)
            print("layer %d d loss / d ln W(%d): backward %.6e, central FD %.6e, rel err %.2e" %
                  (l, mols[k], float(lw.grad[l, k]), fd, err))
            assert err < 1e-5
