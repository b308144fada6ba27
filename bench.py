#!/usr/bin/env python3
"""Benchmark of the MI355X VIE operator on one GPU.

Workload (default `c2`): PWL VIE at 128 MHz on a 64x64x80 grid of 4 mm
voxels holding an ellipsoid head phantom (semi-axes 80/95/110 mm) with 12
seeded piecewise-constant tissue regions (eps_r ~ U[40,80], sigma ~
U[0.3,0.9] S/m, seed 1002); x-polarised plane-wave excitation.  The Green
tensor is assembled on the device (far/medium Gauss rules) plus the near
field on the host; synthetic data, no files.

A step is one application of the full VIE operator (mvp_vie: scatter, N
operator, invG, contrast, gather) in FP64 on one right-hand side, with the
vectors already resident in HBM.  `value` is steps per second.  The operator
runs on the scatterer's support box (its bounding box, FFT lengths >= 2w)
when that is smaller than the grid; the same matvec on the full grid is
timed beside it (`fp64_full_grid`).  The same run also reports the FP32
operator, GMRES solves to 1e-6 (plain FP64 as is_gmres_vie.m, and
FP32-operator mixed precision with FP64 refinement), the HBM roofline of the
operator and the CPU oracle timed on this host.

  python bench.py [--steps K] [--warmup W] [--config c1|c2|c4|c4pwl|c5] [--rhs B] [--no-cpu]
                  [--no-gmres] [--no-batch] [--batch-rhs B] [--no-adjoint] [--no-c5] [--copy-calib]
                  [--no-full-grid]

Sections beside the operator line (all on by default): GMRES solves, the
multi-rhs torso batch (--no-batch), the adjoint gradient at the C3 workload
(--no-adjoint), one 16-rhs group of the body-scale C5 workload (--no-c5), and
the CPU baseline (--no-cpu).  --config c5 times the operator itself on the
128x128x160 body grid, --config c4 / c4pwl on the 96x96x112 torso with the
PWC / PWL basis (--rhs B: B rhs per launch).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "marie-tools_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "c1": {"workload": "pwc_vie_sphere_32x32x32_5mm_298MHz", "basis": "PWC", "freq": 298e6},
    "c2": {"workload": "pwl_vie_head_64x64x80_4mm_128MHz", "basis": "PWL", "freq": 128e6},
    "c4": {"workload": "pwc_vie_torso_96x96x112_3mm_298MHz", "basis": "PWC", "freq": 298e6},
    "c4pwl": {"workload": "pwl_vie_torso_96x96x112_3mm_298MHz", "basis": "PWL", "freq": 298e6},
    "c5": {"workload": "pwl_vie_body_128x128x160_4mm_123MHz", "basis": "PWL", "freq": 123e6},
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(name):
    import numpy as np
    from marie_vie import VieOperator, green, phantoms
    cfg = CONFIGS[name]
    if name == "c1":
        ph = phantoms.sphere()
    elif name in ("c4", "c4pwl"):
        ph = phantoms.layered_torso()
    elif name == "c5":
        ph = phantoms.body_stack()
    else:
        ph = phantoms.ellipsoid_regions()
    pwl = cfg["basis"] == "PWL"
    const = phantoms.constants(cfg["freq"])
    op = VieOperator(ph.dims, cfg["basis"], ph.res)
    t0 = time.perf_counter()
    A = green.assemble(op, "N", const["k0"])
    op.set_green_dense("N", A)
    Mr, Mc = ph.material(cfg["freq"])
    op.set_material(Mr, Mc, ph.idxS)
    t_setup = time.perf_counter() - t0
    ql = 12 if pwl else 3
    Ue = phantoms.project(ph, phantoms.plane_wave(const["k0"]), pwl)
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    b = const["ce"] * np.tile(mcr, ql) * Ue          # ie_solver_vie.m:18
    return {"cfg": cfg, "ph": ph, "op": op, "A": A, "b": b, "Mr": Mr, "Mc": Mc, "pwl": pwl,
            "ql": ql, "k0": const["k0"], "t_setup": t_setup}


def time_matvecs(op, x, precision, steps, warmup):
    import torch
    y = torch.empty_like(x)
    for _ in range(warmup):
        op.apply("VIE", x, precision, out=y)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        op.apply("VIE", x, precision, out=y)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, e0.elapsed_time(e1) / 1e3 / steps, y


def full_grid_matvec(case, xb, y_ref, steps, warmup):
    """The FP64 VIE matvec of the case with the support box switched off
    (VIE_BOX=0: every pass on the full grid) and its difference from the
    box result."""
    import torch
    from marie_vie import VieOperator
    ph = case["ph"]
    os.environ["VIE_BOX"] = "0"
    try:
        op = VieOperator(ph.dims, case["cfg"]["basis"], ph.res)
    finally:
        del os.environ["VIE_BOX"]
    op.set_green_dense("N", case["A"])
    op.set_material(case["Mr"], case["Mc"], ph.idxS)
    op.use_torch_stream()
    wall, ev, y = time_matvecs(op, xb, "fp64", steps, warmup)
    op.close()
    d = float(torch.linalg.vector_norm(y - y_ref) / torch.linalg.vector_norm(y_ref))
    return {"s_per_matvec": ev, "matvec_per_s": 1.0 / wall, "rel_diff_box_vs_full": d}


def cpu_baseline(case, seconds, nmin=5):
    """The oracle (reference-structured numpy restatement) on this host's
    cores, and the parity of the HIP operator with it on the same input.

    Timed: the Tucker T-struct at tol 1e-7 re-expanded on every matvec, as
    mvp_N_pwl_tucker.m does (median of at least nmin matvecs, more while the
    sample stays under `seconds`), and for context the same matvec with the
    dense spectrum precomputed once (median of nmin).  FFTs on scipy.fft with
    one worker per core, tensor products on the BLAS threads.

    Parity: a second device context is fed the SAME T-struct; its matvec on
    the same vector must equal the oracle's to rounding (the default context
    holds the untruncated kernel, which differs from the T-struct by the
    HOSVD truncation; that difference is reported beside it)."""
    import statistics
    import numpy as np
    from marie_vie import VieOperator
    from oracle import vie_oracle as vo
    ncores = len(os.sched_getaffinity(0))
    vo.use_fft_workers(ncores)
    pwl = case["pwl"]
    ph = case["ph"]
    Aop = case["A"] if pwl else case["A"][:, :, :, 0, :]
    t0 = time.perf_counter()
    T = (vo.fft_circ_tucker_pwl if pwl else vo.fft_circ_tucker_pwc)(Aop, 1e-7)
    t_tucker = time.perf_counter() - t0
    log(f"[bench] cpu: HOSVD T-struct {t_tucker:.1f} s")
    isol = vo.idx_solution(ph.idxS, int(np.prod(ph.dims)), case["ql"])
    x = case["b"]
    times = []
    out = None
    t_all = time.perf_counter()
    while len(times) < nmin or (time.perf_counter() - t_all < seconds and len(times) < 3 * nmin):
        t0 = time.perf_counter()
        out = vo.mvp_vie(x, case["Mr"], case["Mc"], isol, ph.dims, ph.res, T, pwl)
        times.append(time.perf_counter() - t0)
        log(f"[bench] cpu: Tucker matvec {len(times)}: {times[-1]:.2f} s")
    med = statistics.median(times)
    t0 = time.perf_counter()
    Td = vo.circulant_dense(Aop, pwl)
    t_dense_setup = time.perf_counter() - t0
    times_d = []
    out_d = None
    for _ in range(nmin):
        t0 = time.perf_counter()
        out_d = vo.mvp_vie(x, case["Mr"], case["Mc"], isol, ph.dims, ph.res, Td, pwl)
        times_d.append(time.perf_counter() - t0)
    med_d = statistics.median(times_d)
    del Td
    log(f"[bench] cpu: median {med:.2f} s/matvec (Tucker), {med_d:.2f} s/matvec (dense spectrum)")
    # the HIP operator fed the same T-struct (support box rebuilt from its terms)
    op = VieOperator(ph.dims, case["cfg"]["basis"], ph.res)
    t0 = time.perf_counter()
    op.set_green_tucker("N", T)
    op.set_material(case["Mr"], case["Mc"], ph.idxS)
    t_gpu_tucker = time.perf_counter() - t0
    y_t = op.apply("VIE", x)
    op.close()
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    return {"value": 1.0 / med, "unit": "matvec/s", "cores": ncores, "kind": "port",
            "sample": f"{len(times)} FP64 VIE matvecs on the full grid, Tucker spectrum (tol 1e-7) "
                      f"re-expanded per matvec as mvp_N_pwl_tucker.m does (median); HOSVD set-up "
                      f"{t_tucker:.1f} s not counted",
            "s_per_matvec": med, "s_per_matvec_all": times,
            "dense_spectrum_s_per_matvec": med_d, "dense_spectrum_setup_s": t_dense_setup,
            "gpu_vs_cpu_relerr": rel(y_t, out),
            "gpu_vs_cpu_relerr_note": "max-abs relative difference, HIP VIE operator (FP64, support box) "
                                      "and oracle both fed the same T-struct",
            "gpu_tucker_setup_s": t_gpu_tucker,
            "cpu_tucker_vs_dense_relerr": rel(out, out_d)}, out_d


def accuracy_vs_converged(op, ph, b0, xm0, x64_0, e_inc0, ce, xg0=None):
    """Both 1e-6 solutions of one rhs (mixed batched, plain FP64) against an
    FP64 solve of the same rhs converged to 1e-10, as currents and as the E
    field recovered from them (em_efield_vie_excitation.m:18-24): each 1e-6
    solution carries its own error of that size, so the mixed one is measured
    against the converged solution, with the plain FP64 one beside it."""
    import numpy as np
    import torch
    from marie_vie import fields
    xr, infr = op.gmres(b0, tol=1e-10, restart=50, maxit=200, precision="fp64")
    xr0 = xr[0]
    nr = torch.linalg.vector_norm(xr0)
    out = {"converged_tol": 1e-10, "converged_matvecs": infr["nmatvec"][0],
           "mixed_x_relerr": float(torch.linalg.vector_norm(xm0 - xr0) / nr),
           "fp64_x_relerr": float(torch.linalg.vector_norm(x64_0 - xr0) / nr)}
    sols = [("ref", xr0), ("mixed", xm0), ("fp64", x64_0)] + ([("mixed_gmres", xg0)] if xg0 is not None else [])
    e = {k: fields.efield(op, ph.idxS, v.cpu().numpy(), e_inc0, ce) for k, v in sols}
    ne = np.linalg.norm(e["ref"])
    out["mixed_E_relerr"] = float(np.linalg.norm(e["mixed"] - e["ref"]) / ne)
    out["fp64_E_relerr"] = float(np.linalg.norm(e["fp64"] - e["ref"]) / ne)
    if xg0 is not None:
        out["mixed_gmres_x_relerr"] = float(torch.linalg.vector_norm(xg0 - xr0) / nr)
        out["mixed_gmres_E_relerr"] = float(np.linalg.norm(e["mixed_gmres"] - e["ref"]) / ne)
    log(f"[bench] accuracy vs FP64 at 1e-10 ({out['converged_matvecs']} matvecs): current mixed "
        f"{out['mixed_x_relerr']:.2e} fp64 {out['fp64_x_relerr']:.2e}; E mixed {out['mixed_E_relerr']:.2e} "
        f"fp64 {out['fp64_E_relerr']:.2e}")
    return out


def batched_section(nrhs=32):
    """Multi-rhs workload (96x96x112 grid of 3 mm voxels, layered torso, 298 MHz,
    PWC, nrhs seeded Hertzian dipoles on a 300 mm cylinder): plain FP64 GMRES
    one rhs at a time (the reference's port loop, ie_solver_vie.m:16-21) versus
    the batched solvers over all rhs at once: mixed-precision GMRES, mixed
    COCR and FP64 block COCG; "speedup" is that of the fastest."""
    import numpy as np
    import torch
    from marie_vie import VieOperator, green, phantoms
    freq = 298e6
    ph = phantoms.layered_torso()
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWC", ph.res)
    t0 = time.perf_counter()
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(freq)
    op.set_material(Mr, Mc, ph.idxS)
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    srcs = phantoms.dipoles(const["k0"], nrhs, 0.30, 0.15, seed=2004)
    B = np.stack([const["ce"] * np.tile(mcr, 3) * phantoms.project(ph, f, False) for f in srcs])
    t_setup = time.perf_counter() - t0
    op.use_torch_stream()
    b = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    log(f"[bench] batched: dims={ph.dims} nS={len(ph.idxS)} nrhs={nrhs} set-up {t_setup:.1f} s")
    # warm the workspaces of both forms
    op.gmres(b[:1].contiguous(), tol=1e-6, restart=50, maxit=1, precision="fp64")
    op.gmres(b, tol=1e-6, restart=50, maxit=1, precision="mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, its = [], []
    for i in range(nrhs):
        x, info = op.gmres(b[i:i + 1].contiguous(), tol=1e-6, restart=50, maxit=200, precision="fp64")
        xs.append(x)
        its.append(info["nmatvec"][0])
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    x64 = torch.cat(xs)
    t0 = time.perf_counter()
    xg, infg = op.gmres(b, tol=1e-6, restart=50, maxit=200, precision="mixed")
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    log(f"[bench] batched: fp64 sequential {t64:.3f} s, mixed GMRES batched {tg:.3f} s, speedup {t64 / tg:.2f}")
    # the mixed batched solve: COCR on the FP32 operator with FP64 replacements
    op.cocr(b, tol=1e-6, maxit=16, precision="mixed")         # workspaces + graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xm, infm = op.cocr(b, tol=1e-6, maxit=2000, precision="mixed")
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    log(f"[bench] batched: mixed COCR batched {tm:.3f} s, speedup {t64 / tm:.2f} "
        f"({np.mean(infm['nmatvec']):.1f} matvecs/rhs [{min(infm['nmatvec'])}, {max(infm['nmatvec'])}], "
        f"{np.mean(infm['nreplace']):.1f} of them FP64); FP32 schedules {op.schedule('fp32')}")
    # block COCG: all rhs in one block Krylov space, FP64 throughout
    op.block_cocg(b, tol=1e-6, maxit=2, precision="fp64")              # workspaces
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xk, infk = op.block_cocg(b, tol=1e-6, maxit=1000, precision="fp64")
    torch.cuda.synchronize()
    tk = time.perf_counter() - t0
    log(f"[bench] batched: FP64 block COCG {tk:.3f} s, speedup {t64 / tk:.2f} "
        f"({infk['nmatvec'][0]} operator applications per rhs, {infk['nreplace'][0]} FP64 residuals)")
    # for context: the same recurrence in FP64 on one rhs (the share of the solver, not of the precision)
    op.cocr(b[:1].contiguous(), tol=1e-6, maxit=16, precision="fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xc64, infc64 = op.cocr(b[:1].contiguous(), tol=1e-6, maxit=2000, precision="fp64")
    torch.cuda.synchronize()
    tc64 = time.perf_counter() - t0

    def relres_of(X):
        return [float(torch.linalg.vector_norm(b[i] - op.apply("VIE", X[i].contiguous())) /
                      torch.linalg.vector_norm(b[i])) for i in range(nrhs)]

    def err_of(X):
        return [float(torch.linalg.vector_norm(X[i] - x64[i]) / torch.linalg.vector_norm(x64[i]))
                for i in range(nrhs)]

    acc = accuracy_vs_converged(op, ph, b[:1].contiguous(), xm[0], x64[0],
                                phantoms.project(ph, srcs[0], False), const["ce"], xg[0])
    rel_m, err_m, rel_g, err_g = relres_of(xm), err_of(xm), relres_of(xg), err_of(xg)
    rel_k, err_k = relres_of(xk), err_of(xk)
    sched32 = op.schedule("fp32")
    op.close()
    best = "block_cocg_fp64" if tk < tm else "cocr_mixed"
    return {"workload": f"pwc_vie_torso_96x96x112_3mm_298MHz_{nrhs}dipoles", "nrhs": nrhs,
            "fp64_sequential_s": t64, "fp64_matvecs_per_rhs_mean": float(np.mean(its)),
            "fastest_batched_solver": best, "speedup": t64 / min(tk, tm),
            "block_cocg_fp64": {"solver": "block COCG, FP64 operator and blocks, all rhs in one block Krylov "
                                          "space (vie_bcocg)",
                                "batched_s": tk, "speedup": t64 / tk,
                                "operator_applications_per_rhs": infk["nmatvec"][0],
                                "fp64_residuals": infk["nreplace"][0], "true_relres_max": max(rel_k),
                                "vs_fp64_x_relerr_max": max(err_k)},
            "mixed_solver": "COCR, FP32 operator, FP64 recurrence scalars and residual replacements (vie_cocr)",
            "mixed_batched_s": tm, "mixed_speedup": t64 / tm,
            "mixed_matvecs_per_rhs_mean": float(np.mean(infm["nmatvec"])),
            "mixed_matvecs_per_rhs_min_max": [min(infm["nmatvec"]), max(infm["nmatvec"])],
            "mixed_fp64_replacements_per_rhs_mean": float(np.mean(infm["nreplace"])),
            "mixed_true_relres_max": max(rel_m), "mixed_vs_fp64_x_relerr_max": max(err_m),
            "axis_fft_schedules_fp32": sched32,
            "mixed_gmres": {"solver": "FP64 refinement around GMRES(20) with the FP32 operator and an "
                                      "FP32-stored basis (vie_gmres)",
                            "batched_s": tg, "speedup": t64 / tg,
                            "matvecs_per_rhs_mean": float(np.mean(infg["nmatvec"])),
                            "true_relres_max": max(rel_g), "vs_fp64_x_relerr_max": max(err_g)},
            "fp64_cocr_rhs0": {"solve_s": tc64, "matvecs": infc64["nmatvec"][0],
                               "fp64_gmres_rhs0_matvecs": its[0]},
            "rhs0_vs_converged": acc, "setup_s": t_setup}


def adjoint_section(nfd=32):
    """Adjoint gradient at the C3 workload (the C2 head phantom at 298 MHz,
    PWL N; f = integral of |E|^2 over a seeded 8x8x8 voxel ROI inside the
    head): forward + adjoint solve time, the gradient against central finite
    differences (step 1e-4) of FP64 device solves on nfd seeded ROI voxels
    (half for eps_r, half for sigma), and the mixed-precision gradient
    against the FP64 one."""
    import numpy as np
    from marie_vie import VieOperator, green, phantoms
    from marie_vie.adjoint import VieProblem
    freq = 298e6
    ph = phantoms.ellipsoid_regions()
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWL", ph.res)
    t0 = time.perf_counter()
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    t_setup = time.perf_counter() - t0
    e_inc = phantoms.project(ph, phantoms.plane_wave(const["k0"]), True)
    mask = np.zeros(int(np.prod(ph.dims)), bool)
    mask[ph.idxS] = True
    mask = mask.reshape(ph.dims, order="F")
    rng = np.random.default_rng(3003)
    n1, n2, n3 = ph.dims
    while True:                                   # an 8x8x8 cube inside the head
        c0 = [int(rng.integers(0, n - 8)) for n in ph.dims]
        if mask[c0[0]:c0[0] + 8, c0[1]:c0[1] + 8, c0[2]:c0[2] + 8].all():
            break
    I, Jy, K = np.meshgrid(*[np.arange(c, c + 8) for c in c0], indexing="ij")
    roi = np.sort((I + n1 * (Jy + n2 * K)).reshape(-1))
    prob = VieProblem(op, ph.idxS, freq, e_inc, roi, tol=1e-12, precision="fp64")
    eps = ph.eps_r.reshape(-1, order="F")[ph.idxS]
    sig = ph.sigma.reshape(-1, order="F")[ph.idxS]
    t0 = time.perf_counter()
    J, info = prob.solve(eps, sig)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_eps, g_sig, info_adj = prob.gradient(J)
    t_adj = time.perf_counter() - t0
    log(f"[bench] adjoint: forward {t_fwd:.3f} s ({info['nmatvec'][0]} matvecs), adjoint {t_adj:.3f} s "
        f"({info_adj['nmatvec'][0]} matvecs)")
    pos = np.searchsorted(ph.idxS, roi)
    pick = rng.choice(pos, size=nfd, replace=False)
    h = 1e-4
    err = {"eps": 0.0, "sig": 0.0}
    for i, v in enumerate(pick):
        which = "eps" if i < nfd // 2 else "sig"
        fs = []
        for sgn in (1.0, -1.0):
            ep, sp = eps.copy(), sig.copy()
            (ep if which == "eps" else sp)[v] += sgn * h
            Jp, _ = prob.solve(ep, sp)
            fs.append(prob.objective(Jp))
        fd = (fs[0] - fs[1]) / (2 * h)
        g = g_eps[v] if which == "eps" else g_sig[v]
        err[which] = max(err[which], abs(g - fd) / abs(fd))
        log(f"[bench] adjoint: fd point {i + 1}/{nfd} ({which}) adjoint {g:.6e} fd {fd:.6e}")
    log(f"[bench] adjoint: gradient vs FD max rel err eps {err['eps']:.2e} sigma {err['sig']:.2e}")
    # the plain FP64 form at the tolerance of the mixed solves below (the FD
    # reference above solves to 1e-12)
    prob.tol = 1e-8
    t0 = time.perf_counter()
    J8, inf8 = prob.solve(eps, sig)
    _, _, inf8_adj = prob.gradient(J8)
    t_fp64_8 = time.perf_counter() - t0
    log(f"[bench] adjoint: FP64 GMRES forward+adjoint at 1e-8 {t_fp64_8:.3f} s "
        f"({inf8['nmatvec'][0]} + {inf8_adj['nmatvec'][0]} matvecs)")
    prob.precision, prob.tol = "mixed", 1e-8
    t0 = time.perf_counter()
    Jm, _ = prob.solve(eps, sig)
    gme, gms, _ = prob.gradient(Jm)
    t_mixed = time.perf_counter() - t0
    d_eps = float(np.linalg.norm(gme - g_eps) / np.linalg.norm(g_eps))
    d_sig = float(np.linalg.norm(gms - g_sig) / np.linalg.norm(g_sig))
    log(f"[bench] adjoint: mixed forward+adjoint {t_mixed:.3f} s, gradient vs FP64 {d_eps:.2e} / {d_sig:.2e}")
    # the same with the COCR solver (forward VIE, adjoint VIE^H)
    prob.solver = "cocr"
    prob.solve(eps, sig)                                           # workspaces + graphs
    t0 = time.perf_counter()
    Jc, infc = prob.solve(eps, sig)
    gce, gcs, infc_adj = prob.gradient(Jc)
    t_cocr = time.perf_counter() - t0
    c_eps = float(np.linalg.norm(gce - g_eps) / np.linalg.norm(g_eps))
    c_sig = float(np.linalg.norm(gcs - g_sig) / np.linalg.norm(g_sig))
    log(f"[bench] adjoint: mixed COCR forward+adjoint {t_cocr:.3f} s ({infc['nmatvec'][0]} + "
        f"{infc_adj['nmatvec'][0]} matvecs), gradient vs FP64 {c_eps:.2e} / {c_sig:.2e}")
    op.close()
    return {"workload": "pwl_vie_head_64x64x80_4mm_298MHz_roi8cube", "setup_s": t_setup,
            "fp64_forward_s": t_fwd, "fp64_adjoint_gradient_s": t_adj,
            "fp64_matvecs": [info["nmatvec"][0], info_adj["nmatvec"][0]],
            "fd_points": nfd, "fd_step": h, "fd_max_relerr_eps": err["eps"], "fd_max_relerr_sigma": err["sig"],
            "fp64_forward_adjoint_tol1e-8_s": t_fp64_8,
            "fp64_tol1e-8_matvecs": [inf8["nmatvec"][0], inf8_adj["nmatvec"][0]],
            "mixed_forward_adjoint_s": t_mixed, "mixed_vs_fp64_grad_relerr_eps": d_eps,
            "mixed_vs_fp64_grad_relerr_sigma": d_sig,
            "mixed_cocr": {"forward_adjoint_s": t_cocr, "matvecs": [infc["nmatvec"][0], infc_adj["nmatvec"][0]],
                           "grad_vs_fp64_relerr_eps": c_eps, "grad_vs_fp64_relerr_sigma": c_sig,
                           "speedup_vs_fp64_tol1e-8": t_fp64_8 / t_cocr}}


def c5_section(nrhs=16):
    """Body-scale batch (C5 workload: 128x128x160 grid of 4 mm, full-body
    ellipsoid stack with 32 seeded tissue regions, 123 MHz, PWL): one group of
    nrhs seeded random-phase surface-current patterns on an enclosing
    cylinder solved to 1e-6 by the batched solvers (mixed GMRES, mixed COCR,
    FP64 block COCG), against one FP64 GMRES solve per rhs (timed on the
    first rhs and scaled)."""
    import numpy as np
    import torch
    from marie_vie import VieOperator, green, phantoms
    freq = 123e6
    ph = phantoms.body_stack()
    const = phantoms.constants(freq)
    free0 = torch.cuda.mem_get_info()[0]
    op = VieOperator(ph.dims, "PWL", ph.res)
    t0 = time.perf_counter()
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(freq)
    op.set_material(Mr, Mc, ph.idxS)
    t_green = time.perf_counter() - t0
    t0 = time.perf_counter()
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    srcs = phantoms.surface_patterns(const["k0"], nrhs, 0.30, 0.30, 6, seed=2005)
    B = np.stack([const["ce"] * np.tile(mcr, 12) * phantoms.project(ph, f, True, nq=2) for f in srcs])
    t_rhs = time.perf_counter() - t0
    op.use_torch_stream()
    b = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    del B
    log(f"[bench] c5: dims={ph.dims} nS={len(ph.idxS)} unknowns/rhs={12 * len(ph.idxS)} nrhs={nrhs} "
        f"green {t_green:.1f} s rhs {t_rhs:.1f} s")
    tiles = op.autotune("VIE", "fp32", nrhs)
    sched = op.schedule("fp32")
    op.gmres(b, tol=1e-6, restart=50, maxit=1, precision="mixed")        # workspaces + graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xg, infg = op.gmres(b, tol=1e-6, restart=50, maxit=200, precision="mixed")
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    relg = max(float(torch.linalg.vector_norm(b[i] - op.apply("VIE", xg[i].contiguous()))
                     / torch.linalg.vector_norm(b[i])) for i in range(nrhs))
    log(f"[bench] c5: mixed GMRES batched {tg:.2f} s ({np.mean(infg['nmatvec']):.0f} matvecs/rhs), "
        f"max true relres {relg:.2e}")
    op.cocr(b, tol=1e-6, maxit=16, precision="mixed")                   # workspaces + graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xm, infm = op.cocr(b, tol=1e-6, maxit=2000, precision="mixed")
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    used = free0 - torch.cuda.mem_get_info()[0]
    rel = max(float(torch.linalg.vector_norm(b[i] - op.apply("VIE", xm[i].contiguous()))
                    / torch.linalg.vector_norm(b[i])) for i in range(nrhs))
    log(f"[bench] c5: mixed COCR batched {tm:.2f} s ({np.mean(infm['nmatvec']):.0f} matvecs/rhs, "
        f"{np.mean(infm['nreplace']):.1f} FP64), max true relres {rel:.2e}, "
        f"device memory in use {used / 2**30:.1f} GiB")
    op.block_cocg(b, tol=1e-6, maxit=2, precision="fp64")              # workspaces
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xk, infk = op.block_cocg(b, tol=1e-6, maxit=1000, precision="fp64")
    torch.cuda.synchronize()
    tk = time.perf_counter() - t0
    relk = max(float(torch.linalg.vector_norm(b[i] - op.apply("VIE", xk[i].contiguous()))
                     / torch.linalg.vector_norm(b[i])) for i in range(nrhs))
    log(f"[bench] c5: FP64 block COCG {tk:.2f} s ({infk['nmatvec'][0]} operator applications per rhs), "
        f"max true relres {relk:.2e}")
    t0 = time.perf_counter()
    x64, inf64 = op.gmres(b[:1].contiguous(), tol=1e-6, restart=50, maxit=200, precision="fp64")
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    err = float(torch.linalg.vector_norm(xm[0] - x64[0]) / torch.linalg.vector_norm(x64[0]))
    log(f"[bench] c5: fp64 one rhs {t64:.2f} s ({inf64['nmatvec'][0]} matvecs); mixed vs fp64 x {err:.2e}; "
        f"est. speedup for {nrhs} rhs {nrhs * t64 / tm:.2f}")
    acc = accuracy_vs_converged(op, ph, b[:1].contiguous(), xm[0], x64[0],
                                phantoms.project(ph, srcs[0], True, nq=2), const["ce"], xg[0])
    op.close()
    return {"workload": "pwl_vie_body_128x128x160_4mm_123MHz_16patterns", "nrhs": nrhs,
            "unknowns_per_rhs": 12 * len(ph.idxS), "green_setup_s": t_green, "rhs_setup_s": t_rhs,
            "mixed_solver": "COCR, FP32 operator, FP64 recurrence scalars and residual replacements (vie_cocr)",
            "mixed_batched_s": tm, "mixed_matvecs_per_rhs_mean": float(np.mean(infm["nmatvec"])),
            "mixed_fp64_replacements_per_rhs_mean": float(np.mean(infm["nreplace"])),
            "mixed_true_relres_max": rel, "fp64_one_rhs_s": t64, "fp64_matvecs": inf64["nmatvec"][0],
            "mixed_vs_fp64_x_relerr_rhs0": err, "rhs0_vs_converged": acc, "speedup_est": nrhs * t64 / tm,
            "mixed_gmres": {"batched_s": tg, "matvecs_per_rhs_mean": float(np.mean(infg["nmatvec"])),
                            "true_relres_max": relg, "speedup_est": nrhs * t64 / tg},
            "block_cocg_fp64": {"batched_s": tk, "operator_applications_per_rhs": infk["nmatvec"][0],
                                "true_relres_max": relk, "speedup_est": nrhs * t64 / tk},
            "fastest_batched_speedup_est": nrhs * t64 / min(tk, tm),
            "device_memory_GiB": used / 2 ** 30, "axis_tiles_fp32": tiles, "axis_fft_schedules_fp32": sched}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--rhs", type=int, default=1,
                    help="right-hand sides per operator application (one batched launch per pass)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gmres", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-full-grid", action="store_true",
                    help="skip the full-grid (VIE_BOX=0) matvec timed beside the support-box one")
    ap.add_argument("--batch-rhs", type=int, default=32, help="rhs of the batched section")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 16-rhs group of the body-scale workload (128x128x160 PWL)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint-gradient section (C3 workload, FD check on 32 voxels)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one plain 512 MiB device copy (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: sharding the right-hand sides over "
                                   "several GPUs is designed but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    if args.copy_calib:
        src = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log(f"[bench] building case {args.config}")
    case = build_case(args.config)
    op = case["op"]
    op.use_torch_stream()
    ph = case["ph"]
    ql = case["ql"]
    nS = len(ph.idxS)
    log(f"[bench] dims={ph.dims} nS={nS} set-up {case['t_setup']:.2f} s; timing {args.steps} steps")
    nb = max(1, args.rhs)
    xb = torch.from_numpy(np.ascontiguousarray(case["b"])).cuda()
    if nb > 1:
        # nb rhs: the plane-wave rhs with a different phase each (one batched launch per pass)
        ph_ = torch.exp(1j * torch.linspace(0, 3.0, nb, dtype=torch.float64, device="cuda"))[:, None]
        xb = (ph_ * xb[None, :]).contiguous()
    wall64, ev64, y64 = time_matvecs(op, xb, "fp64", args.steps, args.warmup)
    wall32, ev32, y32 = time_matvecs(op, xb, "fp32", args.steps, args.warmup)
    # per rhs from here on (one step applies the operator to nb rhs)
    wall64, ev64, wall32, ev32 = wall64 / nb, ev64 / nb, wall32 / nb, ev32 / nb
    err32 = float(torch.linalg.vector_norm(y32 - y64) / torch.linalg.vector_norm(y64))
    log(f"[bench] fp64 {ev64 * 1e3:.3f} ms/matvec, fp32 {ev32 * 1e3:.3f} ms/matvec per rhs, {nb} rhs per launch "
        f"(rel diff {err32:.2e})")
    # the MEX boundary hands over host arrays: the same matvec with the packed
    # vectors crossing PCIe each way (pageable numpy buffers)
    host_io = None
    if nb == 1:   # (a batched run launches only nb-rhs shapes, so its profiles hold one launch shape)
        xh = case["b"].astype(np.complex128)
        op.apply("VIE", xh)
        t0 = time.perf_counter()
        nh = max(1, min(args.steps, 20))
        for _ in range(nh):
            op.apply("VIE", xh)
        host_io = (time.perf_counter() - t0) / nh
        log(f"[bench] fp64 with host vectors (PCIe both ways) {host_io * 1e3:.3f} ms/matvec")
    tiles = {"fp64": op.autotune("VIE", "fp64", nb), "fp32": op.autotune("VIE", "fp32", nb)}
    schedules = {"fp64": op.schedule("fp64"), "fp32": op.schedule("fp32")}
    log(f"[bench] axis passes (z-fwd, y-fwd, y-inv, z-inv): widths {tiles}, FFT schedules {schedules}")

    # the packed operator runs on the scatterer's support box when that is
    # smaller (op.box()); the same matvec on the full grid is timed beside it
    box = op.box()
    full64 = (full_grid_matvec(case, xb[:1].contiguous() if nb > 1 else xb, y64[:1] if nb > 1 else y64, args.steps,
                               args.warmup) if box and not args.no_full_grid else None)
    if full64:
        log(f"[bench] support box {box}; full grid fp64 {full64['s_per_matvec'] * 1e3:.3f} ms/matvec")
    w = box["extent"] if box else ph.dims
    P = box["fft"] if box else op.fft_dims
    oct_box = math.prod(p // 2 + 1 for p in P)
    oct_grid = math.prod(n + 1 for n in ph.dims)
    nG = 60 if case["pwl"] else 6
    # algorithmic bytes per matvec (packed in/out and identity term, the octant
    # spectrum read once): with the grid's octant (n1+1)(n2+1)(n3+1), the
    # figure `achieved` and `frac` use, and with the octant of the support
    # box's FFT lengths, the spectrum this build actually needs
    b_alg64 = 16 * (2 * ql + 1) * nS + 16 * nG * oct_grid / nb
    b_alg64_box = 16 * (2 * ql + 1) * nS + 16 * nG * oct_box / nb
    b_alg32 = 8 * (2 * ql + 1) * nS + 8 * nG * oct_grid / nb
    # five-pass model: A1 = w1 w2 P3 and A2 = w1 P2 P3 per channel, each written
    # and read twice (z-fwd/y-fwd, y-inv/z-inv; y-fwd/x, x/y-inv), no pruning
    a1 = w[0] * w[1] * P[2]
    a2 = w[0] * P[1] * P[2]
    b_pass64 = 16 * ql * (4 * a1 + 4 * a2) + 16 * 2 * ql * nS + 16 * nG * oct_box / nb
    achieved = b_alg64 / ev64 / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.config}_fp64.txt")
    if nb > 1:
        tfile = os.path.join(ROOT, "profiles", f"traffic_{args.config}_fp64_{nb}rhs.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("bytes_per_matvec"):
                traffic = float(line.split("=")[1]) / nb   # the summary counts one launch (nb rhs)

    gm = None
    if not args.no_gmres:
        gm = {}
        # allocate the Krylov workspace outside the timed solves
        op.gmres(xb, tol=1e-6, restart=50, maxit=1, precision="fp64")
        for prec in ("fp64", "mixed"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, info = op.gmres(xb, tol=1e-6, restart=50, maxit=200, precision=prec)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            r = xb - op.apply("VIE", x)
            relres = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(xb))
            gm[prec] = {"solve_s": dt, "iters": info["iters"][0], "matvecs": info["nmatvec"][0],
                        "true_relres": relres}
            gm[prec + "_x"] = x
            log(f"[bench] gmres {prec}: {dt:.3f} s iters {info['iters'][0]} matvecs {info['nmatvec'][0]} "
                f"relres {relres:.2e}")
        # the complex-symmetric short recurrence (vie_cocr): mixed (FP32 operator,
        # FP64 scalars and residual replacements) and, for context, FP64
        op.cocr(xb, tol=1e-6, maxit=16, precision="mixed")        # workspaces + graphs
        op.cocr(xb, tol=1e-6, maxit=16, precision="fp64")
        for prec in ("mixed", "fp64"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, info = op.cocr(xb, tol=1e-6, maxit=2000, precision=prec)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            r = xb - op.apply("VIE", x)
            relres = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(xb))
            key = "cocr_" + prec
            gm[key] = {"solve_s": dt, "matvecs": info["nmatvec"][0], "fp64_replacements": info["nreplace"][0],
                       "true_relres": relres}
            gm[key + "_x"] = x
            log(f"[bench] cocr {prec}: {dt:.3f} s matvecs {info['nmatvec'][0]} (of which FP64 "
                f"{info['nreplace'][0]}) relres {relres:.2e}")
        xd = gm.pop("fp64_x")
        xm = gm.pop("mixed_x")
        xc = gm.pop("cocr_mixed_x")
        gm.pop("cocr_fp64_x")
        nd = torch.linalg.vector_norm(xd)
        gm["mixed_vs_fp64_x_relerr"] = float(torch.linalg.vector_norm(xm - xd) / nd)
        gm["speedup_mixed_vs_fp64"] = gm["fp64"]["solve_s"] / gm["mixed"]["solve_s"]
        gm["cocr_mixed_vs_fp64_x_relerr"] = float(torch.linalg.vector_norm(xc - xd) / nd)
        gm["speedup_cocr_mixed_vs_fp64"] = gm["fp64"]["solve_s"] / gm["cocr_mixed"]["solve_s"]

    batched = None
    if not args.no_batch and not args.no_gmres:
        op.close()
        batched = batched_section(args.batch_rhs)

    adj = None
    if not args.no_adjoint:
        adj = adjoint_section()
    c5 = None
    if not args.no_c5:
        c5 = c5_section()

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle)")
        cpu, cpu_dense_out = cpu_baseline(case, args.cpu_seconds)
        gpu_out = y64.cpu().numpy()
        cpu["gpu_dense_vs_cpu_dense_relerr"] = float(np.abs(gpu_out - cpu_dense_out).max()
                                                     / np.abs(cpu_dense_out).max())
        log(f"[bench] cpu {cpu['s_per_matvec']:.2f} s/matvec; gpu vs cpu, same T-struct "
            f"{cpu['gpu_vs_cpu_relerr']:.2e}, same dense kernel {cpu['gpu_dense_vs_cpu_dense_relerr']:.2e}")

    line = {
        "metric": "GMRES solve s to 1e-6 residual + matvec HBM GB/s; mixed/batched vs plain FP64",
        "value": 1.0 / wall64,   # matvecs (one rhs each) per second
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall64 * nb * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded phantom, Green tensor assembled in-run, plane-wave rhs)",
        "config": {"workload": case["cfg"]["workload"], "grid": list(ph.dims), "res_m": ph.res,
                   "basis": case["cfg"]["basis"], "freq_hz": case["cfg"]["freq"], "n_scat": nS,
                   "unknowns": ql * nS, "rhs_per_step": nb},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "VIE operator application (5 passes) fp64",
                     "alg_bytes_per_matvec": b_alg64,
                     "alg_bytes_note": "B_alg = 16 (2 ql + 1) N_s + 16 n_G (n1+1)(n2+1)(n3+1)",
                     "alg_bytes_box_per_matvec": b_alg64_box,
                     "achieved_box_alg": b_alg64_box / ev64 / 1e9,
                     "frac_box_alg": b_alg64_box / ev64 / 1e9 / HBM_PEAK_GBS,
                     "pass_model_bytes_per_matvec": b_pass64,
                     "event_s_per_matvec": ev64},
        "support_box": box,
        "fp64_full_grid": full64,
        "fp32_operator": {"s_per_matvec": ev32, "matvec_per_s": 1.0 / wall32,
                          "alg_GBs": b_alg32 / ev32 / 1e9, "rel_diff_vs_fp64": err32},
        "fp64_host_vectors_s_per_matvec": host_io,
        "axis_tiles_zf_yf_yi_zi": tiles,
        "axis_fft_schedules_zf_yf_yi_zi": schedules,
        "gmres": gm,
        "batched": batched,
        "adjoint": adj,
        "c5": c5,
        "setup_s": case["t_setup"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
