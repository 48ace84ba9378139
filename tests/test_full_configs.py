"""Parity at the benchmark workloads' full sizes, on their own data:

- the separable suite on uint16 1024x1024x512 (Poisson(100) + 500 blobs):
  exactly the twelve single-axis launches of one benchmark step (Gaussian
  u16->f32 x N=25 s=3, f32->f32 y N=25 s=3, f32->f32 z N=13 s=1.5; mean,
  minimum, maximum N=7 u16->u16 along x, y, z, each pass reading the
  previous pass's output) with the first-launch autotuner on, each launch
  checked on 512 random full lines against the oracle run on the same
  input lines (the tuning launch and the tuned launch both);
- 256 uneven uint16 volumes (axes U{32..256}, Poisson(200)), one random 3D
  filter each: the batched call equals the per-volume calls bit for bit on
  every job, and the oracle on 40 jobs including the largest volume;
- served chains (Gaussian (2,2,1) -> maximum r=2 -> median r=1) on cubes of
  edge 64..768: each result is the same alone and in the mix, and equals
  the oracle chain on the cubes of edge <= 300;
- a 2 GiB uint16 volume (1024^3) through the register-tier median's 64-bit
  addressed gather, on crops against the oracle.
"""
import numpy as np
import pytest

import clk_oracle as O
from helpers import same_bits_or_value

pytestmark = pytest.mark.gpu


def _lines(vol_t, dim, a, b):
    """Full lines along axis `dim` of a (z, y, x) torch volume at the given
    coordinates of the other two axes, as a numpy (n, 1, L) array with the
    filtered axis as x."""
    if dim == 0:
        t = vol_t[a, b, :]
    elif dim == 1:
        t = vol_t[a, :, b]
    else:
        t = vol_t[:, a, b].T
    return t.contiguous().cpu().numpy()[:, None, :]


def _assert_same(got, want, ctx):
    if not same_bits_or_value(got, want):
        bad = np.argwhere(got != want)
        i = tuple(bad[0])
        raise AssertionError("%s: %d mismatches, first at %s: got %r want %r" % (ctx, len(bad), i, got[i], want[i]))


def test_c2_benchmark_step_full_size(gpu):
    import torch
    import bench
    import clkmi
    D, H, W = 512, 1024, 1024
    src = bench.make_c2_volume(torch, D, H, W)  # the benchmark's volume (uint16 in int16 storage)
    f_a = torch.empty((D, H, W), device="cuda", dtype=torch.float32)
    f_b = torch.empty_like(f_a)
    u_a = torch.empty((D, H, W), device="cuda", dtype=torch.int16)
    u_b = torch.empty_like(u_a)
    bufs = {"src": (src, 3), "fa": (f_a, 4), "fb": (f_b, 4), "ua": (u_a, 3), "ub": (u_b, 3)}

    def im(key):
        t, code = bufs[key]
        return clkmi.Image(t.data_ptr(), code, W, H, D)

    def as_np(a, code):
        return a.view(np.uint16) if code == 3 else a

    # (launch, dim, N, s, src buffer, dst buffer): the benchmark's buffer plan
    plan = [("gaussian_blur_separable_3d", 0, 25, 3.0, "src", "fa"),
            ("gaussian_blur_separable_3d", 1, 25, 3.0, "fa", "fb"),
            ("gaussian_blur_separable_3d", 2, 13, 1.5, "fb", "fa")]
    for k in ("mean_separable_3d", "minimum_separable_3d", "maximum_separable_3d"):
        plan += [(k, 0, 7, 0.0, "src", "ua"), (k, 1, 7, 0.0, "ua", "ub"), (k, 2, 7, 0.0, "ub", "ua")]
    rng = np.random.default_rng(22)
    old = clkmi.get_tuning("autotune")
    clkmi.set_tuning("autotune", 1)
    try:
        for name, dim, N, s, sk, dk in plan:
            sizes = {0: (D, H), 1: (D, W), 2: (H, W)}[dim]
            a = torch.from_numpy(rng.integers(0, sizes[0], 512)).cuda()
            b = torch.from_numpy(rng.integers(0, sizes[1], 512)).cuda()
            s_code, d_code = bufs[sk][1], bufs[dk][1]
            src_lines = as_np(_lines(bufs[sk][0], dim, a, b), s_code)
            want = O.separable(name, src_lines, np.float32 if d_code == 4 else np.uint16, 0, N, s)
            # the first launch of a signature times the candidate shapes (all
            # write dst); the second runs the shape it chose
            for rep in range(2):
                bufs[dk][0].zero_()
                getattr(clkmi, name)(im(dk), im(sk), dim, N, s)
                torch.cuda.synchronize()
                got = as_np(_lines(bufs[dk][0], dim, a, b), d_code)
                _assert_same(got[:, 0, :], want[:, 0, :], "%s dim=%d launch %d" % (name, dim, rep))
    finally:
        clkmi.set_tuning("autotune", old)
    del src, f_a, f_b, u_a, u_b, bufs
    torch.cuda.empty_cache()


def _job_oracle(j, src_np):
    dt = np.float32 if j.dst.dtype == 4 else np.uint16
    if j.kind == "gaussian_blur_3d":
        return O.gaussian_blur_3d(src_np, dt, *j.sigma)
    op = O.OP_MIN if j.kind == "minimum_box_3d" else O.OP_MAX
    return O.box_3d(op, src_np, dt, *j.radius)


def test_c4_full_batch(gpu):
    """All 256 jobs of the uneven-batch workload (volumes up to 256 per axis)."""
    import torch
    import bench
    import clkmi
    jobs, bufs = bench.make_c4_jobs(torch, clkmi)
    assert len(jobs) == 256
    stream = torch.cuda.current_stream()
    outs_b = [bufs[2 * i + 1] for i in range(len(jobs))]
    for t in outs_b:
        t.zero_()
    clkmi.batch_filter_3d(jobs, stream=stream)
    torch.cuda.synchronize()
    batched = [t.clone() for t in outs_b]
    for t in outs_b:
        t.zero_()
    for j in jobs:  # the same jobs as one 3D call per volume
        if j.kind == "gaussian_blur_3d":
            clkmi.gaussian_blur_3d(j.src, j.dst, *j.sigma, stream=stream)
        else:
            getattr(clkmi, j.kind)(j.src, j.dst, *j.radius, stream=stream)
    torch.cuda.synchronize()
    for i, (b, p) in enumerate(zip(batched, outs_b)):
        assert torch.equal(b.view(torch.int32) if b.dtype == torch.float32 else b,
                           p.view(torch.int32) if p.dtype == torch.float32 else p), \
            "job %d %s %s: batched != per-volume" % (i, jobs[i].kind, tuple(b.shape))
    vox = [int(j.src.width * j.src.height * j.src.depth) for j in jobs]
    rng = np.random.default_rng(44)
    pick = sorted(set([int(np.argmax(vox))] + [int(v) for v in rng.choice(len(jobs), 39, replace=False)]))
    for i in pick:
        j = jobs[i]
        src_np = bufs[2 * i].cpu().numpy().view(np.uint16)
        got = batched[i].cpu().numpy()
        if got.dtype == np.int16:
            got = got.view(np.uint16)
        _assert_same(got, _job_oracle(j, src_np), "job %d %s %s" % (i, j.kind, src_np.shape))
    del jobs, bufs, batched, outs_b
    torch.cuda.empty_cache()


def test_c5_served_cubes_alone_and_mixed(gpu):
    """Cubes of edge 64, 150, 300 and 768 (the 768 one with rows wide enough
    for the LDS x pass), data as the served workload: each request served
    alone, then all of them at once twice over (so they share batches);
    alone == mix everywhere, == the oracle chain for edges <= 300."""
    import torch
    import bench
    import clkmi
    from clkmi.serve import DEFAULT_CHAIN, FilterServer, GpuExecutor, Request
    edges = [64, 150, 300, 768]
    vols = []
    for i, L in enumerate(edges):
        t = bench.make_c2_volume(torch, L, L, L, seed=50 + i, nblobs=50, amp=1500.0, sigma=4.0, lam=150.0)
        vols.append((L, t, clkmi.Image(t.data_ptr(), 3, L, L, L)))
    srv = FilterServer(GpuExecutor(torch), budget_bytes=60 * 10**9, max_batch=16, n_streams=3)
    try:
        alone = []
        for L, t, im in vols:
            r = srv.submit(Request(im, DEFAULT_CHAIN))
            assert r.wait(300)
            assert r.error is None, r.error
            alone.append(r.outputs[-1].clone())
            del r
        mix = [srv.submit(Request(im, DEFAULT_CHAIN)) for _ in range(2) for L, t, im in vols]
        assert srv.drain(timeout=300)
        for k, r in enumerate(mix):
            assert r.error is None, r.error
            assert torch.equal(r.outputs[-1], alone[k % len(vols)]), "edge %d: mix != alone" % edges[k % len(vols)]
        del mix
    finally:
        srv.close()
    torch.cuda.synchronize()
    for (L, t, im), a in zip(vols, alone):
        if L > 300:
            continue
        data = t.cpu().numpy().view(np.uint16)
        cur = data
        for st in DEFAULT_CHAIN:
            dt = {3: np.uint16, 4: np.float32}[st.dst_dtype]
            if st.kind == "gaussian_blur_3d":
                cur = O.gaussian_blur_3d(cur, dt, *st.sigma)
            elif st.kind == "maximum_box_3d":
                cur = O.box_3d(O.OP_MAX, cur, dt, *st.radius)
            else:
                cur = O.median_box_3d(cur, dt, *(2 * r + 1 for r in st.radius))
        _assert_same(a.cpu().numpy().view(np.uint16), cur, "edge %d vs oracle" % L)
    del vols, alone
    torch.cuda.empty_cache()


def test_median_register_tier_volume_of_2gib(gpu):
    """A uint16 volume of exactly 2 GiB (1024^3) is beyond the 32-bit byte
    offsets of the register tier's buffer-resource gather and takes its
    64-bit addressed gather instead: the 5x5x5 median (the network) and the
    5x5x1 one, checked on crops at corners, edges and the centre against the
    oracle run on the same crops with their halo."""
    import torch
    import clkmi
    D = H = W = 1024
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    src = torch.randint(-32768, 32768, (D, H, W), dtype=torch.int16, device="cuda", generator=g)
    dst = torch.empty_like(src)
    n = (8, 8, 96)  # output crop (z, y, x)
    corners = ((0, 0, 0), (D - 8, H - 8, W - 96), (500, 511, 700), (0, H - 8, 0), (D - 8, 0, W - 96), (3, 2, 1))
    for Ns in ((5, 5, 5), (5, 5, 1)):
        clkmi.median_box_3d(clkmi.Image(dst.data_ptr(), 3, W, H, D), clkmi.Image(src.data_ptr(), 3, W, H, D), *Ns)
        torch.cuda.synchronize()
        halo = ((Ns[2] - 1) // 2, (Ns[1] - 1) // 2, (Ns[0] - 1) // 2)  # (z, y, x)
        for c in corners:
            lo = [max(c[i] - halo[i], 0) for i in range(3)]
            hi = [min(c[i] + n[i] + halo[i], (D, H, W)[i]) for i in range(3)]
            crop = src[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].contiguous().cpu().numpy().view(np.uint16)
            want = O.median_box_3d(crop, np.uint16, *Ns)
            o = [c[i] - lo[i] for i in range(3)]
            want = want[o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]]
            got = dst[c[0]:c[0] + n[0], c[1]:c[1] + n[1], c[2]:c[2] + n[2]].contiguous().cpu().numpy().view(np.uint16)
            _assert_same(got, want, "median %s at %s" % (Ns, c))
