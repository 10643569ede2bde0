"""GPU source-history transport (dendra_amd.history) against the oracle,
bitwise, and against autograd of the reference's formula on the CPU; the
reference's own cases from tests/networks/test_source_history_kernels.py."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from test_history_ref import (FIXED, POST, SOURCE_POS, VALUES, WEIGHTS, packing,
                              torch_project)

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import history_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    return torch.as_tensor(np.asarray(a), device="cuda", dtype=dtype)


def case(rng, dtype, taps, packed, depth=201, n_src=700, n=20000, width=300):
    w = rng.standard_normal(n).astype(dtype)
    src = rng.integers(0, n_src, n)
    post = rng.integers(0, width, n)
    delays = rng.uniform(0.05, 4.9, n)  # ms, dt 0.025: 2..196 steps
    if packed:
        word, mask = packing(n_src)
        history = np.zeros((depth, int(word.max()) + 1), dtype=np.int64)
        bits = rng.uniform(0, 1, (depth, n_src)) < 0.3
        for s in range(depth):
            history = history_ref.record_packed(history, s, bits[s], word, mask)
        meta = {"word_idx": word[src], "bit_mask": mask[src]}
    else:
        history = rng.uniform(-1, 1, (depth, n_src)).astype(dtype)
        meta = {"source_pos": src}
    return w, src, post, delays, history, meta


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("taps", [1, 2, 3])
@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("shape", [(20000, 300, "uniform"), (40000, 45, "uniform"),
                                   (12000, 70, "skewed"), (100, 300, "uniform")])
def test_projection_is_bitwise_the_oracle(dtype, taps, packed, shape):
    """Fan-ins of 66 and 890 connections per target (a tile's connections
    span one to many staged chunks), one target holding half of all
    connections, mostly empty targets, widths off the tile size."""
    from dendra_amd.history import source_history_project
    from dendra_amd.delays import delay_interpolation

    n, width, kind = shape
    rng = np.random.default_rng(3 * taps + packed + n)
    w, src, post, delays, history, meta = case(rng, dtype, taps, packed, n=n, width=width)
    if kind == "skewed":
        post[: n // 2] = 5
        rng.shuffle(post)
    depth, slot, dt = history.shape[0], 150, 0.025
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    kw = {k: dev(v) for k, v in meta.items()}
    if taps == 1:
        steps = np.maximum(np.rint(delays / dt), 1).astype(np.int64)
        got = source_history_project(dev(history), dev([slot]), dev(w), post_idx=dev(post),
                                     depth=depth, width=width, delay_steps=dev(steps), **kw)
        offsets, frac = steps[:, None], None
    else:
        off_t, frac_t = delay_interpolation(dev(delays, tdt), dt, dtype=tdt, taps=taps,
                                            sigma=0.5)
        got = source_history_project(dev(history), dev([slot]), dev(w), post_idx=dev(post),
                                     depth=depth, width=width, delay_taps=(off_t, frac_t),
                                     taps=taps, **kw)
        off = off_t.cpu().numpy()
        offsets = np.stack([off, off + 1], -1) if taps == 2 else off
        frac = frac_t.cpu().numpy()
    col = meta["word_idx"] if packed else meta["source_pos"]
    want = history_ref.project(history, slot, w, post, width, col=col, offsets=offsets,
                               frac=frac, mask=meta.get("bit_mask"), packed=packed)
    g = got.cpu().numpy()
    assert g.dtype == want.dtype
    assert np.array_equal(g.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("taps", [1, 2, 3])
@pytest.mark.parametrize("packed", [False, True])
def test_projection_steps_reuse_records_until_a_source_changes(taps, packed):
    """Steps with the same weights and delays reuse the packed records (the
    slot advances, the history changes); an in-place weight update, new
    delays or a new plan rebuild them.  Every step is bitwise the oracle."""
    from dendra_amd import history as hmod
    from dendra_amd.delays import delay_interpolation

    rng = np.random.default_rng(90 + 2 * taps + packed)
    n, width, dt = 15000, 120, 0.025
    w, src, post, delays, history, meta = case(rng, np.float32, taps, packed, n=n, width=width)
    depth = history.shape[0]
    kw = {k: dev(v) for k, v in meta.items()}
    col = meta["word_idx"] if packed else meta["source_pos"]
    tw, tpost = dev(w), dev(post)

    def delay_args(d):
        if taps == 1:
            steps = np.maximum(np.rint(d / dt), 1).astype(np.int64)
            return {"delay_steps": dev(steps)}, steps[:, None], None
        off_t, frac_t = delay_interpolation(dev(d), dt, dtype=torch.float32, taps=taps,
                                            sigma=0.5)
        off = off_t.cpu().numpy()
        offsets = np.stack([off, off + 1], -1) if taps == 2 else off
        return {"delay_taps": (off_t, frac_t), "taps": taps}, offsets, frac_t.cpu().numpy()

    dargs, offsets, frac = delay_args(delays)
    hits = []
    for step, slot in enumerate((150, 151, 152, 153)):
        if step == 2:  # an optimizer step on the weights, in place
            tw.mul_(0.5)
            w = w * np.float32(0.5)
        if step == 3:  # new delays
            delays = rng.uniform(0.05, 4.9, n)
            dargs, offsets, frac = delay_args(delays)
        h = history if step == 0 else (
            rng.integers(-(2 ** 62), 2 ** 62, history.shape, dtype=np.int64) if packed
            else rng.uniform(-1, 1, history.shape).astype(np.float32))
        before = hmod._records.get(tw)
        got = hmod.source_history_project(dev(h), dev([slot]), tw, post_idx=tpost, depth=depth,
                                          width=width, **dargs, **kw)
        hits.append(before is not None and hmod._records.get(tw)[2] is before[2])
        want = history_ref.project(h, slot, w, post, width, col=col, offsets=offsets,
                                   frac=frac, mask=meta.get("bit_mask"), packed=packed)
        assert np.array_equal(got.cpu().numpy().view(np.uint8), want.view(np.uint8)), step
    assert hits == [False, True, False, False]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("taps", [1, 2, 3])
@pytest.mark.parametrize("packed", [False, True])
def test_workspace_and_direct_entries_agree_with_the_oracle(dtype, taps, packed):
    """dn_history_project_ws_* (records packed at their CSR positions, the
    positions passed in or computed in the workspace) and
    dn_history_project_* (five gathers per connection) give the oracle's
    payload bit for bit; one target holds half of the 12,000 connections."""
    from dendra_amd import _lib
    from dendra_amd.delays import csr_positions, delay_plan

    rng = np.random.default_rng(50 + 2 * taps + packed)
    n, width, slot = 12000, 70, 123
    w, src, post, _, history, meta = case(rng, dtype, taps, packed, n=n, width=width)
    post[: n // 2] = 5
    rng.shuffle(post)
    depth, n_col = history.shape
    off = rng.integers(0, depth, (n, taps)).astype(np.int32)
    frac = None
    if taps == 2:
        frac = rng.uniform(0, 1, n).astype(dtype)
    elif taps == 3:
        frac = rng.dirichlet([2.0, 3.0, 2.0], n).astype(dtype)
    col = (meta["word_idx"] if packed else meta["source_pos"]).astype(np.int32)
    mask = meta.get("bit_mask")
    want = history_ref.project(history, slot, w, post, width, col=col, offsets=off, frac=frac,
                               mask=mask, packed=packed)
    in_ptr, in_edge, _ = delay_plan(dev(post, torch.int64), width)
    th, tw, toff, tcol, ts = dev(history), dev(w), dev(off), dev(col), dev([slot], torch.int64)
    tf = None if frac is None else dev(frac)
    tm = None if mask is None else dev(mask, torch.int64)
    lib, p, sfx = _lib.load(), _lib.ptr, "f32" if dtype == np.float32 else "f64"
    stream = _lib.stream_handle(th.device)
    args = (p(th), int(packed), depth, n_col, p(ts), taps, p(in_ptr), p(in_edge), p(tcol),
            p(tm), p(toff), p(tw), p(tf))
    direct = torch.empty(width, dtype=tw.dtype, device="cuda")
    assert getattr(lib, f"dn_history_project_{sfx}")(*args, p(direct), width, stream) == 0
    nbytes = lib.dn_history_work_bytes(n, taps, int(packed), int(dtype == np.float64), depth,
                                       n_col)
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    packed_out, own = torch.empty_like(direct), torch.empty_like(direct)
    fn = getattr(lib, f"dn_history_project_ws_{sfx}")
    ws = lambda pos: args[:8] + (pos,) + args[8:]
    pos = p(csr_positions(in_edge))
    # without room for the column-major copy of a floating history
    small = lib.dn_history_work_bytes(n, taps, int(packed), int(dtype == np.float64), depth, 0)
    assert small == nbytes if packed else small < nbytes
    no_copy = torch.empty_like(direct)
    assert fn(*ws(pos), p(packed_out), width, n, p(work), nbytes, 0, stream) == 0
    assert fn(*ws(None), p(own), width, n, p(work), nbytes, 0, stream) == 0
    assert fn(*ws(pos), p(no_copy), width, n, p(work), small, 0, stream) == 0
    assert fn(*ws(pos), p(packed_out), width, n, p(work), small - 1, 0, stream) == -1  # EINVAL
    # reuse = 1: the records of the last call serve a later step (another
    # slot, another history) with the same connections
    rng2 = np.random.default_rng(7)
    if packed:
        history2 = rng2.integers(-(2 ** 62), 2 ** 62, history.shape, dtype=np.int64)
    else:
        history2 = rng2.uniform(0, 1, history.shape).astype(dtype)
    slot2 = 57
    want2 = history_ref.project(history2, slot2, w, post, width, col=col, offsets=off,
                                frac=frac, mask=mask, packed=packed)
    th2, ts2 = dev(history2), dev([slot2], torch.int64)
    args2 = (p(th2),) + args[1:4] + (p(ts2),) + args[5:]
    again = torch.empty_like(direct)
    assert fn(*(args2[:8] + (pos,) + args2[8:]), p(again), width, n, p(work), nbytes, 1,
              stream) == 0
    torch.cuda.synchronize()
    for got in (direct, packed_out, own, no_copy):
        assert np.array_equal(got.cpu().numpy().view(np.uint8), want.view(np.uint8))
    assert np.array_equal(again.cpu().numpy().view(np.uint8), want2.view(np.uint8))


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("taps", [None, 2, 3])
def test_reference_fixture_values_and_gradients(packed, taps):
    """test_delivery_uses_current_weights_delays_and_retains_gradients
    (test_source_history_kernels.py:334-405) on the GPU."""
    from dendra_amd.history import source_history_project, source_history_record

    word, mask = packing(3)
    if packed:
        history = torch.zeros((5, 1), dtype=torch.int64, device="cuda")
        for s in range(5):
            history = source_history_record(history, dev([s]), dev(VALUES[s]),
                                            source_word_idx=dev(word),
                                            source_bit_mask=dev(mask))
        meta = {"word_idx": dev(word[SOURCE_POS]), "bit_mask": dev(mask[SOURCE_POS])}
    else:
        history = dev(VALUES).requires_grad_()
        meta = {"source_pos": dev(SOURCE_POS)}
    weights = dev(WEIGHTS).requires_grad_()
    delays = dev([0.15, 0.23, 0.31, 0.42, 0.1], torch.float64).requires_grad_()
    dt = torch.tensor(0.1, dtype=torch.float64)

    def project(w):
        return source_history_project(history, dev([4]), w, post_idx=dev(POST), depth=5,
                                      width=2, delay_steps=dev(FIXED),
                                      delays=delays if taps else None, dt=dt,
                                      taps=taps or 2, sigma=0.4, **meta)

    result = project(weights)
    # the reference's per-edge construction, on the CPU
    ch = torch.as_tensor(VALUES).requires_grad_(not packed)
    cw = torch.as_tensor(WEIGHTS).requires_grad_()
    cd = torch.tensor([0.15, 0.23, 0.31, 0.42, 0.1], dtype=torch.float64, requires_grad=True)
    vals = []
    for e, s in enumerate(SOURCE_POS.tolist()):
        if taps is None:
            gate = ch[(4 - FIXED[e]) % 5, s]
        else:
            dist = (cd[e] / dt).clamp_min(1)
            base = dist.floor()
            if taps == 2:
                fr = dist - base
                gate = ch[(4 - base.long()) % 5, s] * (1 - fr) + \
                    ch[(3 - base.long()) % 5, s] * fr
            else:
                centers = torch.stack((base - 1, base, base + 1))
                interp = torch.softmax(-0.5 * ((dist - centers) / 0.400001) ** 2, dim=0)
                gate = (ch[(4 - centers.long()) % 5, s] * interp).sum()
        vals.append(cw[e] * gate)
    ref = torch.zeros(2, dtype=torch.float64).index_add(0, torch.as_tensor(POST),
                                                        torch.stack(vals))
    torch.testing.assert_close(result.cpu(), ref.detach())
    gpu_t = (weights, *(() if packed else (history,)), *((delays,) if taps else ()))
    cpu_t = (cw, *(() if packed else (ch,)), *((cd,) if taps else ()))
    got = torch.autograd.grad(result.square().sum(), gpu_t)
    exp = torch.autograd.grad(ref.square().sum(), cpu_t)
    for a, b in zip(got, exp, strict=True):
        torch.testing.assert_close(a.cpu(), b)
    torch.testing.assert_close(project(weights * 3).cpu(), (result * 3).detach().cpu())


@pytest.mark.parametrize("taps", [1, 2, 3])
def test_projection_gradients_match_cpu_autograd(taps):
    from dendra_amd.history import source_history_project
    from dendra_amd.delays import delay_interpolation

    rng = np.random.default_rng(40 + taps)
    w, src, post, delays, history, _ = case(rng, np.float64, taps, False, depth=31,
                                            n_src=40, n=600, width=25)
    delays = rng.uniform(0.1, 0.6, 600)
    dt, slot = 0.025, 17
    cot = rng.standard_normal(25)
    steps = np.maximum(np.rint(delays / dt), 1).astype(np.int64)

    def run(device, fn):
        args = [torch.as_tensor(v, device=device).requires_grad_() for v in (history, w, delays)]
        out = fn(*args)
        return out, torch.autograd.grad((out * torch.as_tensor(cot, device=device)).sum(),
                                        args, allow_unused=True)

    def gpu(h, tw, td):
        kw = dict(delay_steps=dev(steps)) if taps == 1 else dict(delays=td, dt=dt, taps=taps,
                                                                sigma=0.5)
        return source_history_project(h, dev([slot]), tw, post_idx=dev(post), depth=31,
                                      width=25, source_pos=dev(src), **kw)

    def cpu(h, tw, td):
        if taps == 1:
            return torch_project(h, slot, tw, post, 25, src, steps[:, None])
        off, fr = delay_interpolation(td, dt, dtype=torch.float64, taps=taps, sigma=0.5)
        offsets = torch.stack([off, off + 1], -1) if taps == 2 else off
        depth = h.shape[0]
        rows = (slot - offsets).remainder(depth)
        col = torch.as_tensor(src)
        if taps == 2:
            values = tw * (h[rows[:, 0], col] * (1.0 - fr) + h[rows[:, 1], col] * fr)
        else:
            g = torch.stack([h[rows[:, j], col] for j in range(3)], -1)
            values = tw * (g * fr).sum(-1)
        return torch.zeros(25, dtype=torch.float64).index_add(0, torch.as_tensor(post), values)

    og, gg = run("cuda", gpu)
    oc, gc = run("cpu", cpu)
    if taps < 3:  # one and two taps: the same correctly rounded divisions
        assert torch.equal(og.detach().cpu(), oc.detach())
    else:         # Gaussian tap weights: exp on two devices
        torch.testing.assert_close(og.detach().cpu(), oc.detach(), rtol=1e-13, atol=1e-14)
    for a, b in zip(gg, gc):
        if b is None:
            assert a is None or not a.any()
            continue
        torch.testing.assert_close(a.cpu(), b, rtol=1e-12, atol=1e-13)


def test_record_float_and_packed():
    from dendra_amd.history import read_packed_history, source_history_record

    rng = np.random.default_rng(5)
    size, depth = 130, 4
    word, mask = packing(size)
    history = torch.full((depth, 3), 17, dtype=torch.int64, device="cuda")
    gates = np.zeros(size)
    active = [0, 61, 62, 63, 64, 125, 126, 129]
    gates[active] = 1
    out = source_history_record(history, dev([2]), dev(gates), source_word_idx=dev(word),
                                source_bit_mask=dev(mask))
    want = history_ref.record_packed(np.full((depth, 3), 17), 2, gates, word, mask)
    assert np.array_equal(out.cpu().numpy(), want)
    assert (history == 17).all()
    seen = read_packed_history(out, dev([2]).expand(size), dev(word), dev(mask),
                               dtype=torch.float64)
    assert np.array_equal(seen.cpu().numpy(), gates)
    owned = history.clone()
    ptr = owned.data_ptr()
    donated = source_history_record(owned, dev([2]), dev(gates), source_word_idx=dev(word),
                                    source_bit_mask=dev(mask), _inplace=True)
    assert donated.data_ptr() == ptr and torch.equal(donated, out)
    # floating: only the unwritten history keeps its gradient
    h = torch.arange(15, dtype=torch.float64, device="cuda").reshape(5, 3).requires_grad_()
    g = dev([-0.2, 0.0, 0.8], torch.float64).requires_grad_()
    for own in ({}, {"_clone_once": True}):
        upd = source_history_record(h, dev([7]), g, **own)
        exp = history_ref.record_float(np.arange(15.0).reshape(5, 3), 7, [-0.2, 0.0, 0.8])
        assert np.array_equal(upd.detach().cpu().numpy(), exp)
        gh, gg = torch.autograd.grad(upd.sum(), (h, g))
        m = torch.ones(5, 3, dtype=torch.float64)
        m[2] = 0
        assert torch.equal(gh.cpu(), m) and torch.equal(gg.cpu(), torch.ones(3, dtype=torch.float64))
    big = rng.standard_normal((201, 5000)).astype(np.float32)
    gates32 = rng.standard_normal(5000).astype(np.float32)
    upd = source_history_record(dev(big), dev([205]), dev(gates32))
    assert np.array_equal(upd.cpu().numpy(), history_ref.record_float(big, 205, gates32))


def test_empty_and_donated_outputs():
    from dendra_amd.history import source_history_project, source_history_record

    empty = torch.empty(0, dtype=torch.long, device="cuda")
    for packed in (False, True):
        history = torch.zeros((3, 1 if packed else 0),
                              dtype=torch.int64 if packed else torch.float64, device="cuda")
        meta = {"word_idx": empty, "bit_mask": empty} if packed else {"source_pos": empty}
        out = source_history_project(history, dev([2]),
                                     torch.empty(0, dtype=torch.float64, device="cuda"),
                                     post_idx=empty, depth=3, width=4, delay_steps=empty, **meta)
        assert torch.equal(out.cpu(), torch.zeros(4, dtype=torch.float64))
        upd = source_history_record(history, dev([2]),
                                    torch.empty(0, dtype=torch.float64, device="cuda"),
                                    source_word_idx=empty if packed else None,
                                    source_bit_mask=empty if packed else None)
        assert torch.equal(upd, history)
    scratch = torch.full((2,), 99.0, dtype=torch.float64, device="cuda")
    res = source_history_project(dev(VALUES), dev([4]), dev(WEIGHTS), post_idx=dev(POST),
                                 depth=5, width=2, delay_steps=dev(FIXED),
                                 source_pos=dev(SOURCE_POS), _scratch=scratch)
    assert res is scratch
    want = history_ref.project(VALUES, 4, WEIGHTS, POST, 2, col=SOURCE_POS,
                               offsets=FIXED[:, None])
    assert np.array_equal(res.cpu().numpy(), want)
    with pytest.raises(ValueError, match=r"source_pos entries must lie in \[0, 3\)"):
        source_history_project(dev(VALUES), dev([4]), dev(WEIGHTS), post_idx=dev(POST),
                               depth=5, width=2, delay_steps=dev(FIXED),
                               source_pos=dev([0, 1, 2, 3, 0]))
