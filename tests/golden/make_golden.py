"""Writes tests/golden/reference_outputs.npz: outputs of the reference's own CPU code
(oracle/_ref/libmkoref.so, built from the reference sources by oracle/Makefile) on small seeded
inputs of this repository's workloads, so the oracle stays pinned where the reference build is
absent. Run from the repository root: python tests/golden/make_golden.py"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from mikroptim_amd import abi, systems  # noqa: E402
from oracle_lib import reference  # noqa: E402


def main():
    ref = reference()
    if ref is None:
        raise SystemExit("the reference build oracle/_ref/libmkoref.so is missing")
    out = {}
    b, v = systems.singlet(4, seed=7)
    for name, fail in (("soft", abi.FAILURE_SOFT), ("hard", abi.FAILURE_HARD)):
        s = systems.singlet_settings(fail)
        res, opd, st = ref.residuals(b, s)
        m, f, h = ref.merit(b, s)
        x0 = v.extract(b)[:2]
        J, r0, ch = ref.fd_jacobian(b.take([0]), s, v, x0)
        Ja, r0a, sta = ref.adjoint_jacobian(b.take([0]), s, v, x0)
        out.update({f"singlet_{name}_residuals": res, f"singlet_{name}_opd": opd,
                    f"singlet_{name}_status": st, f"singlet_{name}_merit": m,
                    f"singlet_{name}_fd_J": J, f"singlet_{name}_fd_r0": r0,
                    f"singlet_{name}_adj_J": Ja, f"singlet_{name}_x0": x0})
    ob, ov = systems.objective(variants=2, seed=2)
    s = systems.objective_settings(abi.FAILURE_HARD)
    s.rings, s.spokes = 3, 4
    res, opd, st = ref.residuals(ob, s)
    out.update({"objective_coarse_residuals": res, "objective_coarse_status": st})
    out["objective_paraxial_v1"] = ref.paraxial(ob, 1)
    # constraint rows (ExtraResiduals / TraceExtraResiduals) on the singlet
    s = systems.singlet_settings(abi.FAILURE_HARD)
    out["singlet_extrema"] = ref.surface_extrema(b, s)
    s.constraints = systems.singlet_constraints()
    out["singlet_rows"] = ref.constraint_rows(b, s)
    out["singlet_rows_merit"] = ref.merit(b, s)[0]
    soft = systems.singlet_settings(abi.FAILURE_SOFT)
    soft.constraints = systems.singlet_constraints()
    x0 = v.extract(b)[:2]
    T = ref.residuals(b.take([0]), systems.singlet_settings(abi.FAILURE_SOFT))[0].shape[1]
    for key, fn in (("fd", ref.fd_jacobian), ("adjoint", ref.adjoint_jacobian)):
        J, r0, _ = fn(b.take([0, 1]), soft, v, x0)
        out[f"singlet_{key}_rows_J"] = J[:, :, T:]
        out[f"singlet_{key}_rows_r0"] = r0[:, T:]
    # sine-condition and distortion rows, and their FD Jacobian columns
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_image_constraints()
    out["singlet_image_rows"] = ref.constraint_rows(b, s)
    J, _, _ = ref.fd_jacobian(b.take([0, 1]), s, v, v.extract(b)[:2])
    out["singlet_image_fd_J"] = J[:, :, J.shape[2] - len(s.constraints):]
    s = systems.singlet_settings(abi.FAILURE_HARD)
    s.constraints = systems.singlet_constraints()
    # FP32 prefilter stats (prefilter_stats_cpu) of the multistart singlet batch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_prefilter import stats  # noqa: E402
    mb, _ = systems.singlet_multistart(150, seed=11)
    st = stats(ref, mb, systems.singlet_settings(abi.FAILURE_HARD))
    for name in ("merit", "optical_length", "failed", "uncertain_ok", "uncertain_failed"):
        out[f"prefilter_multistart_{name}"] = getattr(st, name)
    lb, lv = systems.singlet(8, seed=3)
    for key, method in (("fd", abi.JACOBIAN_FD), ("adjoint", abi.JACOBIAN_ADJOINT)):
        opt = abi.lm_options(max_iterations=8, threads=2, jacobian_method=method)
        x, _, _ = ref.optimize_lm(lb.take([0]), s, lv, lv.extract(lb)[:2], opt)
        out[f"singlet_rows_lm_{key}_x"] = x
    # Buchdahl fits (fit_buchdahl) of the catalogue glasses: n0, nu1, nu2, nu3, max, rms per
    # glass, for 2 / 3 dispersion terms and Chebyshev / equidistant nodes
    from mikroptim_amd import system_file as sf
    for cat in ("basis", "synthetisch"):
        glasses = sf.load_glass_catalog(os.path.join(ROOT, "tests", "golden", "data", "glass",
                                                     cat + ".json"))
        for terms in (2, 3):
            for nodes in ("chebyshev", "equidistant"):
                rows = [ref.fit_buchdahl(g, terms, nodes=nodes) for g in glasses.values()
                        if g.range_nm[0] <= 400.0 and g.range_nm[1] >= 700.0]
                out[f"buchdahl_fit_{cat}_{terms}_{nodes}"] = np.array(rows)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "reference_outputs.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
