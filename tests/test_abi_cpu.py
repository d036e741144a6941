"""C-ABI library on the CPU: loads, exports every declared symbol, honours the
host-side contract (empty shapes, no device, plan validation, planner tiers,
cache file), and the LDS staging layout passes its known-answer and
bank-conflict checks."""
import os
import re

import numpy as np
import pytest

import maxmulsk_amd as mm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "maxmulsk.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(maxmulsk_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = mm.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(mm.EXPORTED)


def test_status_values_match_reference_enum():
    # capi/maxmulsk.h:24-28
    assert (mm.OK, mm.ERR_UNSUPPORTED, mm.ERR_ALLOC) == (0, 1, 2)


def test_empty_shapes_contract():
    # capi/maxmulsk.cpp:13-20: M or N == 0 -> OK, nothing touched
    C = np.full((3, 5), 7.0, np.float32)
    A = np.ones((3, 4), np.float32)
    B = np.ones((4, 5), np.float32)
    for fn in (mm.sgemm_v3, mm.sgemm_v4):
        C[:] = 7.0
        assert fn(A, B, C, 0, 4, 5) == mm.OK
        assert fn(A, B, C, 3, 4, 0) == mm.OK
        assert np.all(C == 7.0)
        # K == 0 -> C cleared
        C[:] = 7.0
        assert fn(A, B, C, 3, 0, 5) == mm.OK
        assert np.all(C == 0.0)
    # K == 0 with null A/B is fine: they are never read
    C[:] = 7.0
    assert mm.sgemm_v4(None, None, C, 3, 0, 5) == mm.OK and np.all(C == 0.0)


@pytest.mark.skipif(os.environ.get("MAXMULSK_EXPECT_GPU") == "1", reason="GPU present")
def test_no_device_is_an_error_not_a_crash():
    if mm.device_count() > 0:
        pytest.skip("a gfx950 device is present")
    A = np.ones((8, 8), np.float32)
    C = np.full((8, 8), 3.0, np.float32)
    for fn in (mm.sgemm_v3, mm.sgemm_v4):
        assert fn(A, A, C, 8, 8, 8) == mm.ERR_UNSUPPORTED
        assert np.all(C == 3.0)
        assert "gfx950" in mm.last_error()


def test_null_pointers_rejected():
    C = np.zeros((2, 2), np.float32)
    assert mm.sgemm_v4(None, None, C, 2, 2, 2) == mm.ERR_UNSUPPORTED
    assert mm.sgemm_v4(None, None, None, 2, 2, 2) == mm.ERR_UNSUPPORTED


def test_pack_api_rejects_bad_arguments_without_a_device_call():
    lib = mm.lib()
    h = mm.C.c_void_p(1234)
    X = np.ones(64, np.float32)
    assert lib.maxmulsk_pack(X.ctypes.data, 8, 8, 0, 1, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    assert h.value is None  # out is cleared on failure
    assert lib.maxmulsk_pack(None, 8, 8, 0, 1, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    assert lib.maxmulsk_pack(X.ctypes.data, 8, 8, 2, 1, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    assert lib.maxmulsk_pack(X.ctypes.data, 8, 8, 0, 5, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    assert lib.maxmulsk_pack(X.ctypes.data, 0, 8, 0, 1, None, mm.C.byref(h)) == mm.ERR_UNSUPPORTED
    assert lib.maxmulsk_sgemm_packed(None, None, X.ctypes.data, None) == mm.ERR_UNSUPPORTED
    lib.maxmulsk_packed_destroy(None)


def test_phase_counters_names_and_reset():
    mm.reset_phase_counters()
    c = mm.phase_counters()
    assert list(c) == ["allocation", "stage_in", "pack", "compute", "reduce", "writeback", "zero_fill"]
    assert all(v == 0 for v in c.values())
    assert mm.lib().maxmulsk_phase_counter_name(99) is None
    assert mm.lib().maxmulsk_phase_counter(-1) == 0


def test_invalid_plan_rejected_before_anything_runs():
    A = np.ones((8, 8), np.float32)
    C = np.full((8, 8), 9.0, np.float32)
    for plan in [(32, 1), (100, 1), (64, 0), (64, 33), (256, -1)]:
        assert mm.sgemm_ex(A, A, C, 8, 8, 8, plan=plan) == mm.ERR_UNSUPPORTED
        assert np.all(C == 9.0)
        assert mm.workspace_bytes(8, 8, 8, 0, plan) == 0
    assert mm.sgemm_ex(A, A, C, 8, 8, 8, mode=7) == mm.ERR_UNSUPPORTED


def test_candidates_and_rule_plan():
    cands = mm.candidates()
    assert len(cands) == 18
    from maxmulsk_amd import selector as sel
    assert sel.NCAND == len(cands) == mm.lib().maxmulsk_num_candidates()
    assert cands == [(t, s) for t in (64, 128, 256) for s in (1, 2, 4, 8, 16, 32)]
    assert mm.rule_plan(4096, 4096, 4096) == (256, 1)
    # narrow M: small tiles, K split to fill the CUs (8 tiles x 32 slices of 256)
    t, s = mm.rule_plan(64, 8192, 512)
    assert t == 64 and s == 32
    t, s = mm.rule_plan(64, 512, 32768)
    assert s >= 1 and t in (64, 128, 256)
    for shape in [(1, 1, 1), (1023, 999, 1025), (67, 8191, 513), (4096, 1536, 24576)]:
        t, s = mm.rule_plan(*shape)
        assert (t, s) in cands


def _rule_split(K):
    return min(32, max(1, -(-K // 4096)))


def test_grouped_batch_plans():
    """Default plans of a grouped launch (pure host logic, no device): a small
    launch gets the per-problem rule (tile edge from N, ceil(K/4096) slices);
    a launch of at least 1024 one-slice tiles runs unsliced, except a problem
    one of whose tiles holds more than the launch's work per CU."""
    none = lambda M, K, N: (None, None, None, M, K, N)
    small = [none(37, 1536, 2112), none(300, 999, 1025), none(512, 16384, 512), none(5, 0, 7), none(0, 9, 9),
             none(200, 64, 96)]
    for mode in (0, 1, 2):
        got = mm.grouped_plans(small, mode)
        for (_, _, _, M, K, N), p in zip(small, got):
            if M == 0 or N == 0 or K == 0:
                assert p == (0, 0)
            else:
                assert p == mm.grouped_default_plan(M, K, N, mode) == (256 if N >= 256 else 128 if N >= 128 else 64,
                                                                       _rule_split(K))
    # the 256-problem workload: 8510 one-slice tiles, no tile longer than the work per CU
    rng = np.random.default_rng(2024)
    c4 = [none(int(rng.integers(1, 513)), int(rng.choice([512, 1536, 7168, 16384])), int(rng.choice([512, 2112, 4096, 7168])))
          for _ in range(256)]
    for mode in (0, 1):
        assert all(p == (256, 1) for p in mm.grouped_plans(c4, mode))
    # 600 short problems (1200 tiles) and one whose two tiles are each longer than the work per CU
    mixed = [none(64, 8192, 512)] * 600 + [none(64, 1 << 20, 512)]
    got = mm.grouped_plans(mixed, 1)
    assert got[:600] == [(256, 1)] * 600 and got[600] == (256, 32)
    assert mm.grouped_plans([], 0) == []


def test_workspace_bytes():
    assert mm.workspace_bytes(64, 8192, 512, 0, (64, 16)) == 16 * 64 * 512 * 4
    # 256-wide tiles on more than 128 rows take the packed path: the workspace
    # holds A and B rewritten as fragments (FP32 copy, or 2 / 3 BF16 terms),
    # padded to whole 256-row tiles and whole packed k-blocks (32 deep in plain
    # and split3, 16 in split6) ...
    assert mm.workspace_bytes(4096, 4096, 4096, 0, (256, 1)) == 2 * 4096 * 4096 * 4
    assert mm.workspace_bytes(4095, 4096, 4096, 1, (256, 1)) == 2 * 2 * 4096 * 4096 * 2
    assert mm.workspace_bytes(4096, 4096, 4000, 2, (256, 1)) == 2 * 3 * 4096 * 4096 * 2
    assert mm.workspace_bytes(4096, 4090, 4096, 2, (256, 1)) == 2 * 3 * 4096 * 4096 * 2  # K % 16 != 0
    # ... except in the split modes on whole tiles and k-blocks, where the
    # direct 256 x 256 tile splits on the fly: split-K partials only
    assert mm.workspace_bytes(4096, 4096, 4096, 1, (256, 1)) == 0
    assert mm.workspace_bytes(4096, 4096, 4096, 2, (256, 1)) == 0
    assert mm.workspace_bytes(4096, 4096, 4096, 1, (256, 2)) == 2 * 4096 * 4096 * 4
    assert mm.workspace_bytes(512, 4080, 768, 2, (256, 4)) == 4 * 512 * 768 * 4
    assert mm.workspace_bytes(300, 20, 300, 1, (256, 1)) == 2 * (2 * 512 * 32 * 2)
    assert mm.workspace_bytes(300, 40, 300, 0, (256, 1)) == 2 * (512 * 64 * 4)
    assert mm.workspace_bytes(300, 40, 300, 1, (256, 1)) == 2 * (2 * 512 * 64 * 2)
    assert mm.workspace_bytes(300, 40, 300, 2, (256, 1)) == 2 * (3 * 512 * 48 * 2)
    assert mm.workspace_bytes(4096, 4096, 4096, 0, (128, 1)) == 0
    # split clamps to the number of k-blocks: K=64 with BK=32 has 2 blocks
    assert mm.workspace_bytes(64, 64, 64, 0, (64, 16)) == 2 * 64 * 64 * 4
    assert mm.workspace_bytes(0, 64, 64, 0, (64, 2)) == 0


def test_planner_tiers(monkeypatch):
    lib = mm.lib()
    lib.maxmulsk_cache_clear()
    monkeypatch.delenv("MAXMULSK_PLAN", raising=False)
    plan, src = mm.choose_plan(300, 700, 500)
    assert src == 3 and plan == mm.rule_plan(300, 700, 500)
    assert lib.maxmulsk_cache_put(300, 700, 500, 0, mm.Plan(128, 4)) == mm.OK
    plan, src = mm.choose_plan(300, 700, 500)
    assert src == 1 and plan == (128, 4)
    # other modes are separate keys
    assert mm.choose_plan(300, 700, 500, 1)[1] == 3
    monkeypatch.setenv("MAXMULSK_PLAN", "64,2")
    assert mm.choose_plan(300, 700, 500) == ((64, 2), 0)
    for bad in ["64", "64,2,3", "63,2", "64,99", "x,y", ""]:
        monkeypatch.setenv("MAXMULSK_PLAN", bad)
        assert mm.choose_plan(300, 700, 500)[1] == 1  # malformed override ignored
    lib.maxmulsk_cache_clear()
    assert lib.maxmulsk_cache_put(1, 1, 1, 0, mm.Plan(65, 1)) == mm.ERR_UNSUPPORTED


def test_cache_file_round_trip(tmp_path):
    lib = mm.lib()
    lib.maxmulsk_cache_clear()
    entries = {(64, 8192, 512, 0): (64, 16), (4096, 4096, 4096, 1): (256, 1), (7, 9, 11, 2): (128, 2)}
    for (M, K, N, mode), p in entries.items():
        assert lib.maxmulsk_cache_put(M, K, N, mode, mm.Plan(*p)) == mm.OK
    path = str(tmp_path / "tune.bin").encode()
    assert lib.maxmulsk_cache_save(path) == mm.OK
    lib.maxmulsk_cache_clear()
    assert lib.maxmulsk_cache_size() == 0
    assert lib.maxmulsk_cache_load(path) == mm.OK
    assert lib.maxmulsk_cache_size() == 3
    for (M, K, N, mode), p in entries.items():
        assert mm.choose_plan(M, K, N, mode) == (p, 1)
    # a corrupted or foreign file is refused and leaves the cache unchanged
    raw = bytearray(open(path, "rb").read())
    raw[8] = 99  # version
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(raw))
    assert lib.maxmulsk_cache_load(str(bad).encode()) == mm.ERR_UNSUPPORTED
    raw = bytearray(open(path, "rb").read())
    raw[16:22] = b"gfx942"
    bad.write_bytes(bytes(raw))
    assert lib.maxmulsk_cache_load(str(bad).encode()) == mm.ERR_UNSUPPORTED
    assert lib.maxmulsk_cache_load(b"/nonexistent/x.bin") == mm.ERR_DEVICE
    lib.maxmulsk_cache_clear()


# ---- LDS layout ------------------------------------------------------------------

# (mode, rows, bk) for every operand tile the kernels stage: tile edges 64/128/256,
# plain and split3 at bk 32, split6 at bk 32 and 16 (256x256 tiles).
LAYOUTS = [(m, r, bk) for m in (0, 1, 2) for r in (64, 128, 256) for bk in (32,)] + [(2, 256, 16)]
LU = {0: 4, 1: 8, 2: 8}
TERMS = {0: 1, 1: 2, 2: 3}


def offsets(mode, op, rows, bk):
    cpr = bk // LU[mode]
    out = np.full((TERMS[mode], rows, cpr), -1, np.int64)
    for t in range(TERMS[mode]):
        for r in range(rows):
            for c in range(cpr):
                out[t, r, c] = mm.lds_chunk_offset(mode, op, rows, bk, r, c, t)
    return out


@pytest.mark.parametrize("mode,rows,bk", LAYOUTS)
@pytest.mark.parametrize("op", [0, 1])
def test_lds_layout_is_a_bijection_onto_the_stage(mode, rows, bk, op):
    off = offsets(mode, op, rows, bk)
    assert off.min() >= 0
    total = TERMS[mode] * rows * bk * (4 if mode == 0 else 2)
    assert off.max() + 16 <= total
    assert np.all(off % 16 == 0)
    assert len(np.unique(off)) == off.size == total // 16


@pytest.mark.parametrize("mode,rows,bk", LAYOUTS)
def test_lds_layout_known_answer(mode, rows, bk):
    """A[i][k] = 1000*i + k staged into an LDS image, then read back through the
    MFMA fragment pattern: lane l of fragment (block, k-group) must see row
    block*32 + l%32 and k = kgroup*2*LU + (l//32)*LU + e (test_sme.cpp:265-273)."""
    lu = LU[mode]
    cpr = bk // lu
    op = 0
    img = np.full(TERMS[mode] * rows * bk * 2, -999.0)  # one slot per element
    off = offsets(mode, op, rows, bk)
    for t in range(TERMS[mode]):
        for r in range(rows):
            for c in range(cpr):
                base = off[t, r, c] // 16 * lu
                for e in range(lu):
                    img[base + e] = 1000 * r + c * lu + e + 1e6 * t
    kgroups = cpr // 2
    for t in range(TERMS[mode]):
        for blk in range(rows // 32):
            for kg in range(kgroups):
                for lane in range(64):
                    r = blk * 32 + lane % 32
                    c = 2 * kg + lane // 32
                    base = mm.lds_chunk_offset(mode, op, rows, bk, r, c, t) // 16 * lu
                    got = img[base:base + lu]
                    want = [1000 * r + kg * 2 * lu + (lane // 32) * lu + e + 1e6 * t for e in range(lu)]
                    assert list(got) == want


def _worst_conflict(offs, lanes_per_cycle, bytes_per_cycle):
    worst = 1
    for g in range(0, 64, lanes_per_cycle):
        slots = {}
        for o in offs[g:g + lanes_per_cycle]:
            if o is None:
                continue
            s = (o % bytes_per_cycle) // 16
            slots[s] = slots.get(s, 0) + 1
        if slots:
            worst = max(worst, max(slots.values()))
    return worst


def _geometries():
    return [(g["cfg"], g["mode"]) for g in mm.tile_infos()]


def _check_tile_lds_patterns(mode, bm, bn, bk, nt, bw):
    lu = LU[mode]
    cpr = bk // lu
    insts = []
    # A writer: unit u = tid + nt*i -> (row u // cpr, chunk u % cpr)
    total = bm * cpr
    for i in range((total + nt - 1) // nt):
        for w in range(nt // 64):
            lanes = []
            for l in range(64):
                u = w * 64 + l + nt * i
                lanes.append(None if u >= total else mm.lds_chunk_offset(mode, 0, bm, bk, u // cpr, u % cpr, 0))
            insts.append(lanes)
    # B writer: unit u -> (nb = u % (bn/bw), chunk = u // (bn/bw)); store j writes column bw*nb+j
    nu = bn // bw
    total = cpr * nu
    for i in range((total + nt - 1) // nt):
        for j in range(bw):
            for w in range(nt // 64):
                lanes = []
                for l in range(64):
                    u = w * 64 + l + nt * i
                    lanes.append(None if u >= total else
                                 mm.lds_chunk_offset(mode, 1, bn, bk, bw * (u % nu) + j, u // nu, 0))
                insts.append(lanes)
    # readers of both operands
    reads = []
    for op, rows in ((0, bm), (1, bn)):
        for blk in range(rows // 32):
            for kg in range(cpr // 2):
                reads.append([mm.lds_chunk_offset(mode, op, rows, bk, blk * 32 + l % 32, 2 * kg + l // 32, 0)
                              for l in range(64)])
    # ds_write_b128 serves 8 consecutive lanes per 128-byte cycle, ds_read_b128
    # 16 lanes per 256-byte cycle (both measured on gfx950 with strided and
    # grouped lane patterns).
    for lanes in insts:
        assert _worst_conflict(lanes, 8, 128) == 1
    for lanes in reads:
        assert _worst_conflict(lanes, 16, 256) == 1


@pytest.mark.parametrize("cfg,mode", _geometries())
def test_lds_access_patterns_are_bank_conflict_free(cfg, mode):
    """Simulates the tile kernel's writer and reader lane patterns for every
    geometry and mode: every group of 8 lanes of a store (128-byte LDS cycle)
    and of 16 lanes of a fragment read (256-byte cycle) hits distinct 16-byte
    bank groups. This is necessary, not sufficient: the B fragment reads pass
    it, yet cost 4 conflict cycles per ds_read_b128 on the hardware, while the
    A fragment reads and every store measured none."""
    g = mm.tile_info(cfg, mode)
    _check_tile_lds_patterns(mode, g["bm"], g["bn"], g["bk"], g["threads"], g["b_cols"])


@pytest.mark.parametrize("mode,bm,bn,bk,threads,b_cols", [
    # eight-wave tiles of the split modes: direct 128 x 128, grouped 64 x 128 .. 128 x 256
    *[(m, bm, bn, 32, 512, 2) for m in (1, 2) for bm, bn in ((128, 128), (64, 128), (64, 256), (128, 256))],
    (0, 64, 64, 64, 256, 4), (1, 64, 64, 128, 256, 4),     # deep 64 x 64 for small grids
    (1, 64, 64, 64, 256, 4), (1, 128, 128, 64, 512, 2),    # 64-deep 64 x 64 (two per CU), split3 128 x 128
    # plain tiles of the grouped kernel: 16-deep k-blocks
    *[(0, bm, bn, 16, 256, 4) for bm, bn in ((64, 64), (64, 128), (64, 256), (128, 128), (128, 256))],
])
def test_lds_patterns_of_the_launch_variants(mode, bm, bn, bk, threads, b_cols):
    """Layouts launched besides the table's: the eight-wave tiles of the split
    modes (two-column B units, as every eight-wave tile; the direct 128 x 128
    tile and every tile of the grouped kernel), the deep tiles (64 x 64: plain
    k-blocks of 64, split3 of 128 or 64; split3 128 x 128: 64) and the plain
    tiles of the grouped kernel (k-blocks of 16)."""
    _check_tile_lds_patterns(mode, bm, bn, bk, threads, b_cols)


def test_tile_geometries():
    infos = mm.tile_infos()
    assert len(infos) == 18
    for g in infos:
        assert g["threads"] == 64 * g["waves"] and g["waves"] in (4, 8)
        # LDS budget of one CU: two stages of both operands
        per_elem = 4 if g["mode"] == 0 else 2 * {1: 2, 2: 3}[g["mode"]]
        assert 2 * (g["bm"] + g["bn"]) * g["bk"] * per_elem <= 160 * 1024


def test_product_input_generator_matches_oracle():
    from maxmulsk_amd import inputs
    from oracle import oracle as orc
    assert np.array_equal(inputs.uniform(9, (123, 45)), orc.uniform(9, (123, 45)))
    assert np.array_equal(inputs.log_uniform(9, (77,)), orc.log_uniform(9, (77,)))


@pytest.mark.parametrize("mode,rows,bk,threads,b_cols", [
    (0, 64, 64, 256, 4), (1, 64, 128, 256, 4), (1, 64, 64, 256, 2), (2, 64, 64, 256, 2), (2, 64, 32, 256, 1),
    (1, 128, 64, 512, 2), (2, 128, 32, 256, 2)])
def test_lds_patterns_of_the_decoupled_feed(mode, rows, bk, threads, b_cols):
    """The producer waves of the decoupled feed stage the tile (64 x 64, or
    128 x 128 in the split modes) with their own thread indices and the direct
    tile's staging units for that many threads (B columns per unit as for that
    wave count and depth), into ring slots of the direct tile's layout."""
    _check_tile_lds_patterns(mode, rows, rows, bk, threads, b_cols)


@pytest.mark.parametrize("mode,bm,bn,b_cols", [
    (m, bm, bn, 2 if bn == 128 else 4) for m in (1, 2) for bm, bn in ((64, 128), (64, 256), (128, 128), (128, 256))])
def test_lds_patterns_of_the_grouped_feed(mode, bm, bn, b_cols):
    """The grouped launch on the decoupled feed stages the direct grouped
    kernel's four split-mode geometries into 32-deep slots with four producer
    waves (B columns per unit as the direct tile takes for four waves)."""
    assert mm.tile_info(1 if bn == 128 and bm == 64 else (2 if bm == 64 else (3 if bn == 128 else 4)), mode)["bk"] == 32
    _check_tile_lds_patterns(mode, bm, bn, 32, 256, b_cols)


def test_lds_epilogue_of_the_256_tile():
    """The 256 x 256 tile's LDS-staged epilogue (wave tiles 64 wide, the only
    width the hook accepts): each wave's stores land inside its
    32 x (wtn + 4) block without overlap; every element of the 32-row block is
    read back exactly once, by the lane that stores it to the same row and
    column of C as the register epilogue (accumulator value 4q + e of fragment
    column tj in lane l: row l % 32, column tj * 32 + 8q + 4(l / 32) + e);
    stores and reads are free of bank conflicts; eight waves' blocks fit in
    the tile's 128 KiB of LDS."""
    wtn = 64
    block = 32 * (wtn + 4)
    assert 8 * block * 4 <= 128 * 1024
    held = {}
    for tj in range(wtn // 32):
        for q in range(4):
            offs = [mm.lds_epilogue_write(wtn, lane, tj, q) for lane in range(64)]
            assert _worst_conflict([4 * o for o in offs], 8, 128) == 1, (tj, q)
            for lane, o in enumerate(offs):
                assert 0 <= o and o + 4 <= block
                for e in range(4):
                    assert o + e not in held
                    held[o + e] = (lane % 32, tj * 32 + 8 * q + 4 * (lane // 32) + e)
    assert len(held) == 32 * wtn
    seen = set()
    for i in range(wtn // 8):
        reads = [mm.lds_epilogue_read(wtn, lane, i) for lane in range(64)]
        assert _worst_conflict([4 * o for o, _, _ in reads], 16, 256) == 1, i
        for o, row, col in reads:
            for e in range(4):
                assert held[o + e] == (row, col + e)
                seen.add((row, col + e))
    assert len(seen) == 32 * wtn
    assert mm.lds_epilogue_write(32, 0, 0, 0) == -1 and mm.lds_epilogue_write(wtn, 64, 0, 0) == -1
    assert mm.lds_epilogue_read(wtn, 0, wtn // 8)[0] == -1


def test_feed_variant_rule():
    """Which kernel a launch takes under each feed setting (pure host code):
    the automatic rule feeds launches of at most one workgroup per CU: 64 x 64
    split6 with K slices of at least 256, 64 x 64 split3 with slices of at
    least 1024, 128 x 128 split3 and split6 whose tiles and k-blocks are all
    full (slices of at least 1024), and 128 x 128 split3 with M, N >= 256
    and slices of at least 2048; forcing the feed picks a variant for
    every 64 x 64 launch and every split-mode 128 x 128 launch; other
    geometries never take it."""
    prev = mm.get_feed()
    try:
        mm.set_feed(mm.FEED_AUTO)
        auto = {
            (1024, 1024, 1024, (64, 1)): (0, 2, 3),
            (768, 768, 4096, (64, 1)): (0, 1, 3),
            (640, 640, 3000, (64, 1)): (0, 2, 3),
            (512, 768, 4500, (64, 1)): (0, 1, 3),
            (512, 384, 2500, (64, 1)): (0, 2, 3),           # 48 workgroups
            (64, 512, 8192, (64, 32)): (0, 0, 3),           # slices of 256
            (64, 512, 8192, (64, 64)): (-1, -1, -1),        # split above 32
            (64, 512, 4096, (64, 32)): (0, 0, 0),           # slices of 128
            (64, 512, 8192, (64, 16)): (0, 0, 3),           # split3: slices of 512
            (67, 513, 8191, (64, 16)): (0, 0, 0),           # 288 workgroups
            (4096, 4096, 4096, (256, 1)): (0, 0, 0),
            (1024, 1024, 1024, (128, 1)): (0, 7, 8),        # 128 x 128, full tiles and k-blocks
            (2048, 2048, 2048, (128, 1)): (0, 7, 8),
            (1920, 1920, 2000, (128, 1)): (0, 0, 0),        # partial k-block, slices < 2048
            (1000, 1000, 4000, (128, 1)): (0, 7, 0),        # ragged, slices >= 2048
            (1000, 1100, 3000, (128, 1)): (0, 7, 0),
            (1000, 1101, 3000, (128, 1)): (0, 7, 0),        # odd N
            (200, 1100, 3000, (128, 1)): (0, 0, 0),         # M < 256
            (2048, 2048, 512, (128, 1)): (0, 0, 0),         # slices of 512
            (4096, 4096, 1024, (128, 1)): (0, 0, 0),        # 1024 workgroups
        }
        for (M, N, K, plan), want in auto.items():
            assert tuple(mm.feed_variant(M, K, N, mode, plan) for mode in (0, 1, 2)) == want, (M, N, K, plan)
        mm.set_feed(mm.FEED_DECOUPLED)
        assert [mm.feed_variant(67, 8191, 513, m, (64, 16)) for m in (0, 1, 2)] == [4, 5, 6]
        assert [mm.feed_variant(768, 4096, 768, m, (64, 1)) for m in (0, 1, 2)] == [4, 1, 3]
        assert [mm.feed_variant(64, 8192, 512, m, (64, 32)) for m in (0, 1, 2)] == [4, 2, 3]
        assert mm.feed_variant(1024, 1024, 1024, 1, (256, 1)) == 0
        assert [mm.feed_variant(1000, 3000, 1100, m, (128, 1)) for m in (0, 1, 2)] == [0, 7, 8]
        assert [mm.feed_variant(4096, 1024, 4096, m, (128, 1)) for m in (0, 1, 2)] == [0, 7, 8]
        assert mm.feed_variant(64, 64, 64, 1, (48, 1)) == -1
        mm.set_feed(mm.FEED_DIRECT)
        assert mm.feed_variant(1024, 1024, 1024, 1, (64, 1)) == 0
        with pytest.raises(ValueError):
            mm.set_feed(-1)
        assert mm.get_feed() == mm.FEED_DIRECT
    finally:
        mm.set_feed(prev)


def test_stream_probe_rejects_bad_arguments_before_any_device_call():
    """maxmulsk_probe_stream refuses the packed plan, an invalid plan, K or N not
    a multiple of 4 and a non-positive call count with MAXMULSK_ERR_UNSUPPORTED,
    before it touches a device; without a device a valid request is refused
    the same way instead of crashing."""
    import ctypes as C
    lib = mm.lib()
    us = (C.c_double * 3)()
    bad = [(256, 256, 256, (256, 1), 10), (64, 512, 512, (48, 1), 10), (64, 255, 512, (64, 1), 10),
           (64, 512, 255, (64, 1), 10), (64, 512, 512, (64, 1), 0), (64, 512, 512, (64, 33), 10)]
    for M, K, N, plan, calls in bad:
        st = lib.maxmulsk_probe_stream(M, K, N, C.byref(mm.Plan(*plan)), calls, None, us)
        assert st == mm.ERR_UNSUPPORTED, (M, K, N, plan, calls)
    if mm.device_count() == 0:
        st = lib.maxmulsk_probe_stream(64, 512, 512, C.byref(mm.Plan(64, 1)), 10, None, us)
        assert st == mm.ERR_UNSUPPORTED
        assert "device" in mm.last_error()
