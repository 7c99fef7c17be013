"""Benchmark: D3(BJ) SPME energy + forces + stress throughput on one MI355X.

Workload: the 100k-atom H/C/N/O cubic box (ratio 2 : 1 : 0.3 : 0.7,
1.0 g/cm^3, 1.0 A minimum separation, seed 0), r_cut = 6 A, SPME with 5
B-spline nodes, 1.2 A mesh spacing (mesh 256^3), PBE BJ parameters,
c6tol = 1 %.  One step = one full evaluation through the HIP library: cell
list + CN, CN -> C6 charges, spreading, 3-D FFT, reciprocal-space kernel,
inverse FFT, gathering with the adjoint C6/CN chain -> energy, forces and
the 3x3 stress.  Positions are resident in HBM before the timed region.

Prints one JSON line; see --help.  Extra legs: the per-kernel roofline of the
dominant kernel (HIP events on the launch stream, recorded over the timed
steps), the FP64 CPU oracle on a bounded sample (cpu_baseline) and the
reference algorithm written with plain torch ops on the GPU (plain_gpu).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "E+F+stress atom-evals/s, SPME D3 ~100k atoms; batched+adjoint vs plain GPU"
HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def _args():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--atoms", type=int, default=100_000)
    ap.add_argument("--c6tol", type=float, default=1.0)
    ap.add_argument("--cnfunc", default="smooth_cut", choices=["smooth_cut", "d4"],
                    help="CN model (d4: k-space D4 CN, no neighbour list)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-gpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel HIP events")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for HBM counters)")
    return ap.parse_args()


def kernel_bytes(name: str, plan: dict) -> float | None:
    """Compulsory HBM bytes of one launch of `name` (algorithmic, not measured):
    M = mesh points, Mh = half-spectrum points, C = channels, N = atoms."""
    C, M, Mh, N = plan["C"], plan["M"], plan["Mh"], plan["N"]
    table = {
        "k_r2c_z": 4 * M * C + 8 * Mh * C,       # read real mesh, write half spectrum
        "k_c2c_y_fwd": 16 * Mh * C,              # read + write complex half spectrum
        "k_kspace": 16 * Mh * C,                 # read S, write dE/drho spectrum (x passes fused)
        "k_c2c_y_inv": 16 * Mh * C,
        "k_c2r_z": 8 * Mh * C + 4 * M * C,
        "k_fill": 4 * M * C,                     # zero the real mesh
        "k_spread": 4 * M * C,                   # mesh written once (accumulated in L2)
        # spreading fused into the z pass: the real mesh never reaches HBM
        "k_spread_r2c_z": 8 * Mh * C + 36 * N,   # write half spectrum; read positions, species, charges
        "k_c2r_z_gather": 8 * Mh * C + 36 * N,   # read half spectrum; per-atom records in, accumulators out
    }
    return table.get(name)


def run_gpu(args):
    import torch

    from fourierd3_amd import FourierD3ASECalculator
    from fourierd3_amd import _lib as L
    from fourierd3_amd import workloads as W

    dev = torch.device("cuda:0")
    box = W.hcno_box(args.atoms, seed=0)
    calc = FourierD3ASECalculator(device=dev, c6tol=args.c6tol, cnfunc=args.cnfunc)
    from types import SimpleNamespace
    calc._build_model(SimpleNamespace(numbers=box.numbers, cell=box.cell, pbc=np.array([True] * 3)))
    model = calc._model
    pos = torch.tensor(box.positions, dtype=torch.float64, device=dev)
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"energy": torch.empty(1, **f64), "grad_pos": torch.empty((box.n_atoms, 3), **f64),
           "virial": torch.empty((3, 3), **f64), "grad_params": torch.empty(4, **f64)}
    for _ in range(args.warmup):
        calc.evaluate(pos, box.cell, out=out)
    torch.cuda.synchronize()
    ns = [int(v) for v in model.ns_mesh]
    M = ns[0] * ns[1] * ns[2]
    plan = {"C": model.n_species * model.n_rank, "M": M, "Mh": ns[0] * ns[1] * (ns[2] // 2 + 1), "N": box.n_atoms}
    def instrumented(steps):
        # HIP events on the launch stream around every kernel
        L.profile_only(None)
        L.profile_enable(True)
        for _ in range(steps):
            calc.evaluate(pos, box.cell, out=out)
        torch.cuda.synchronize()
        rep = L.profile_report()
        L.profile_enable(False)
        return rep

    prof = {}
    prof_steps = min(args.steps, 10)
    roof_name = None
    if not args.no_profile:
        # the roofline kernel: the longest one with an algorithmic byte count
        pick = instrumented(2)
        roof_name = max((k for k in pick if kernel_bytes(k, plan)), key=lambda k: pick[k][0], default=None)
    # timed steps; only the roofline kernel carries an event pair per launch
    if roof_name:
        L.profile_only(roof_name)
        L.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        calc.evaluate(pos, box.cell, out=out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = (t1 - t0) / args.steps
    timed = {}
    if roof_name:
        timed = L.profile_report()
        L.profile_enable(False)
        L.profile_only(None)
    # per-kernel times from separate, instrumented steps after the timed ones
    if not args.no_profile:
        prof = instrumented(prof_steps)
    energy_ev = float(out["energy"].item()) * 27.21138505
    kernels = {k: {"ms_per_launch": v[0] / v[1], "launches_per_step": v[1] / prof_steps} for k, v in prof.items()}
    roof = None
    if roof_name and roof_name in timed:
        name = roof_name
        avg_s = timed[name][0] / timed[name][1] / 1e3
        b = kernel_bytes(name, plan)
        achieved = b / avg_s / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                traffic = json.load(fh).get(f"{name}@{args.atoms}")
        roof = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved * 1e9 / HBM_PEAK, "traffic": traffic, "algorithmic_bytes": b,
                "avg_ms": avg_s * 1e3, "launches_timed": timed[name][1]}
    info = {"box": box, "model": model, "energy_ev": energy_ev, "dt": dt, "kernels": kernels, "roofline": roof,
            "plan": plan}
    return info


def plain_gpu(box, model, steps=3):
    """The reference algorithm written with plain torch ops on the GPU in
    float32 (rfftn, index_add_, complex bmm, autograd through a strain tensor),
    as the reference module runs it; edges from the HIP neighbour list, built
    outside the timed region as the reference's benchmark does."""
    import ctypes as C
    import math

    import torch

    from fourierd3_amd import _lib as L

    dev = torch.device("cuda:0")
    f32 = torch.float32
    n = box.n_atoms
    A2B = 1.0 / 0.52917726
    p = torch.tensor(box.positions, dtype=torch.float64, device=dev)
    arr = L.cell_array(box.cell)
    nb = C.c_size_t()
    L.check(L.lib().fd3_neighbours_workspace(n, arr, 6.0, C.byref(nb)))
    ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    L.check(L.lib().fd3_neighbours_count(n, L.ptr(p), 1, arr, 6.0, L.ptr(ws), ws.numel(), L.ptr(cnt), stream))
    e = int(cnt.item())
    src = torch.empty(e, dtype=torch.int64, device=dev)
    dst = torch.empty_like(src)
    sh = torch.empty((e, 3), dtype=torch.int32, device=dev)
    L.check(L.lib().fd3_neighbours_fill(n, L.ptr(p), 1, arr, 6.0, L.ptr(ws), ws.numel(), e, L.ptr(src), L.ptr(dst),
                                        L.ptr(sh), stream))
    unit_shifts = sh.to(f32)
    spec = model.species
    rcov = model.rcov.to(f32)
    cnref = model.cnref.to(f32)
    vq = model.v_q_reshaped.to(f32)
    eigs = model.eigs.to(f32)
    qq = model._t_qq
    prm = [float(v) for v in model.params.tolist()]
    R = prm[2] * torch.sqrt(3 * qq) + prm[3]
    selfc = prm[0] / torch.diagonal(R) ** 6 + 3 * prm[1] * torch.diagonal(qq) / torch.diagonal(R) ** 8
    nx, ny, nz = (int(v) for v in model.ns_mesh)
    order = 5
    rc_b = 6.0 * A2B
    mx = torch.fft.fftfreq(nx, d=1.0 / nx, device=dev)
    my = torch.fft.fftfreq(ny, d=1.0 / ny, device=dev)
    mz = torch.fft.rfftfreq(nz, d=1.0 / nz, device=dev)
    sc = (torch.sinc(mx / nx)[:, None, None] * torch.sinc(my / ny)[None, :, None]
          * torch.sinc(mz / nz)[None, None, :]) ** (2 * order)
    sc = torch.where(sc < 1e-10, 1e-10, sc)
    wz = torch.full((nz // 2 + 1,), 2.0, device=dev)
    wz[0] = 1.0
    wz[-1] = 1.0
    nS, r = model.n_species, model.n_rank
    Rv = R.view(nS, nS, 1, 1, 1)
    Qv = qq.view(nS, nS, 1, 1, 1)
    pi = math.pi
    s8, c8 = math.sin(pi / 8), math.cos(pi / 8)

    def weights5(x):
        x2, x3, x4 = x * x, x ** 3, x ** 4
        return torch.stack([(1 - 2 * x) ** 4 / 384, (19 - 44 * x + 24 * x2 + 16 * x3 - 16 * x4) / 96,
                            (115 - 120 * x2 + 48 * x4) / 192, (19 + 44 * x + 24 * x2 - 16 * x3 - 16 * x4) / 96,
                            (1 + 2 * x) ** 4 / 384])

    sa, sb, sc3 = torch.meshgrid(torch.arange(5, device=dev), torch.arange(5, device=dev),
                                 torch.arange(5, device=dev), indexing="ij")
    sa, sb, sc3 = sa.flatten(), sb.flatten(), sc3.flatten()

    def step():
        cell = torch.tensor(box.cell, dtype=f32, device=dev)
        strain = torch.zeros(3, 3, dtype=f32, device=dev, requires_grad=True)
        scell = cell + torch.einsum("ab,Ab->Aa", strain, cell)
        cb = scell * A2B
        vol = torch.abs(torch.linalg.det(cb))
        pos = torch.tensor(box.positions, dtype=f32, device=dev, requires_grad=True)
        spos = pos + torch.einsum("ab,ib->ia", strain, pos)
        pb = spos * A2B
        shb = (unit_shifts @ scell) * A2B
        # k-space filter rebuilt on the full mesh every call, as _update_cell does
        recip = 2 * pi * torch.linalg.inv(cb).T
        kv = (mx.view(-1, 1, 1, 1) * recip[0] + my.view(1, -1, 1, 1) * recip[1]
              + mz.view(1, 1, -1, 1) * recip[2])
        k = torch.linalg.norm(kv, dim=-1).view(1, 1, nx, ny, nz // 2 + 1)
        small = k * Rv < 1e-15
        ks = torch.where(small, 1.0, k)
        x = ks * Rv
        ft6 = torch.where(small, 2 * pi ** 2 / (3 * Rv ** 3) - 2 * pi ** 2 * Rv / 9 * k ** 2,
                          2 * pi ** 2 / (3 * Rv ** 4) / ks * (torch.exp(-x) - 2 * torch.exp(-x / 2)
                                                              * torch.cos(pi / 3 + x * math.sqrt(3) / 2)))
        c1 = pi ** 2 * math.sqrt(2) * s8 / Rv ** 5
        ft8 = torch.where(small, c1 - c1 * Rv ** 2 / 6 * k ** 2,
                          -pi ** 2 / Rv ** 6 / ks * (torch.exp(-x * s8) * torch.cos(pi / 4 + x * c8)
                                                     + torch.exp(-x * c8) * torch.cos(3 * pi / 4 + x * s8)))
        kf = (prm[0] * ft6 + Qv * 3 * prm[1] * ft8) / sc * wz
        kfilter = kf.flatten(2).to(torch.complex64).permute(2, 0, 1).contiguous()
        # CN
        vec = pb[dst] - pb[src] + shb
        rr = torch.linalg.norm(vec, dim=-1)
        rs = torch.clamp(rr, min=1e-6)
        rcs = rcov[spec[src]] + rcov[spec[dst]]
        r1 = 0.5 * rcs + 0.5 * rc_b
        rcf = torch.where(rs > r1, rs, r1)
        kcn = 16.0 + (rcf - r1) ** 2 / ((rc_b - rcf) ** 2 + 1e-6)
        contrib = torch.sigmoid(kcn * (4.0 / 3.0 * rcs / rs - 1.0))
        contrib = torch.where(rr > 1e-6, contrib, torch.zeros_like(contrib))
        cn = torch.zeros(n, dtype=f32, device=dev).index_add(0, src, contrib)
        refs = cnref[spec]
        logits = (-4.0 * (cn.unsqueeze(1) - refs) ** 2).masked_fill(refs == -1, float("-inf"))
        wts = torch.softmax(logits, dim=1)
        c6 = torch.einsum("nk,nkr->nr", wts, vq[spec])
        onehot = torch.zeros(n, nS, r, dtype=f32, device=dev).index_put((torch.arange(n, device=dev), spec), c6)
        vself = torch.dot(eigs, (c6 ** 2 * selfc[spec].unsqueeze(1)).sum(0))
        # B-spline weights and spreading (torch-pme P3M, 5 nodes)
        u = torch.tensor([nx, ny, nz], dtype=f32, device=dev) * (pb @ torch.linalg.inv(cb))
        base = torch.round(u)
        w = weights5(u - base)                                      # (5, N, 3)
        idx = torch.stack([(base + j) % torch.tensor([nx, ny, nz], device=dev) for j in range(-2, 3)]).long()
        wgeo = (w[sa, :, 0] * w[sb, :, 1] * w[sc3, :, 2]).t()
        lin = (idx[sc3, :, 2] + nz * (idx[sb, :, 1] + ny * idx[sa, :, 0])).t().flatten()
        vals = (onehot.reshape(n, 1, -1) * wgeo.unsqueeze(2)).flatten(0, 1)
        rho = torch.zeros(nx * ny * nz, nS * r, dtype=f32, device=dev).index_add(0, lin, vals)
        rho = rho.t().contiguous().view(nS, r, nx, ny, nz)
        S = torch.fft.rfftn(rho, dim=(2, 3, 4))
        ready = S.flatten(2).permute(2, 0, 1).contiguous()
        interm = torch.bmm(kfilter, ready.conj())
        fh = (ready * interm).real.sum(dim=(0, 1))
        energy = (torch.dot(eigs, fh) / (-2 * vol) + vself / 2) * 27.21138505
        energy.backward()
        return energy.detach(), -pos.grad, strain.grad / vol * A2B ** 3

    out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"value": n / dt, "unit": "atom-evals/s", "ms_per_step": dt * 1e3, "dtype": "f32",
            "energy_ev": float(out[0]), "what": "reference algorithm in plain torch ops on the GPU (FP32, autograd)"}


def cpu_baseline(budget_s: float):
    """FP64 oracle on a bounded sample of the workload: a 12.5k-atom box of the
    same composition, density and separation (1/8 of the atoms, mesh 128^3),
    timed for whole calculator calls (neighbour list + E + F + stress)."""
    import torch

    from fourierd3_amd import workloads as W
    from oracle.fd3_oracle import OracleFourierD3
    from oracle.neighbours import neighbour_list_kdtree

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    torch.set_num_threads(cores)
    box = W.hcno_box(12_500, seed=0)
    o = OracleFourierD3(box.numbers, box.cell)
    runs, t_all, t_nl = 0, 0.0, 0.0
    while runs < 1 or (t_all < budget_s and runs < 5):
        t0 = time.perf_counter()
        edges = neighbour_list_kdtree(box.positions, box.cell, 6.0)
        t1 = time.perf_counter()
        o.calculate(box.positions, box.cell, edges=edges)
        t2 = time.perf_counter()
        t_all += t2 - t0
        t_nl += t1 - t0
        runs += 1
    return {"value": box.n_atoms * runs / t_all, "unit": "atom-evals/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (torch CPU), {runs} full calculator evaluation(s) of a 12.5k-atom H/C/N/O box "
                      f"of the benchmark's composition and density (mesh 128^3), neighbour list included; "
                      f"without it {box.n_atoms * runs / (t_all - t_nl):.4g} atom-evals/s"}


def main():
    args = _args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    info = run_gpu(args)
    model, box = info["model"], info["box"]
    if args.calibrate_copy:
        import torch
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    res = {
        "metric": METRIC,
        "value": box.n_atoms / info["dt"],
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": info["dt"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"H/C/N/O cubic box {box.n_atoms} atoms, SPME E+F+stress",
                   "n_atoms": box.n_atoms, "box_A": float(box.cell[0, 0]), "mesh": [int(v) for v in model.ns_mesh],
                   "species": int(model.n_species), "rank": int(model.n_rank), "c6tol": args.c6tol,
                   "r_cut_A": 6.0, "interpolation_nodes": 5, "mesh_spacing_A": 1.2, "cnfunc": args.cnfunc},
        "energy_eV": info["energy_ev"],
        "roofline": info["roofline"],
        "kernels": info["kernels"],
    }
    if not args.no_plain_gpu:
        try:
            pg = plain_gpu(box, model)
            res["plain_gpu"] = pg
            res["speedup_vs_plain_gpu"] = res["value"] / pg["value"]
        except Exception as exc:  # report, do not fail the benchmark line
            res["plain_gpu"] = {"error": repr(exc)[:300]}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
