// LDS staging layout shared by the GEMM kernels (device) and the host-side
// known-answer tests (maxmulsk_lds_chunk_offset).
//
// Both operand tiles are stored in "fragment-major" order. A fragment is what
// one wave hands to one MFMA as its A or B operand: 64 lanes x 16 bytes = 1 KiB.
// Lane l of a fragment holds row (l % 32) of a 32-row block and the 16-byte
// chunk h = l / 32 of a 2-chunk k-group. A chunk is LU consecutive k values:
// 4 FP32 values (plain mode) or 8 BF16 values (split modes). The MFMA operand
// layouts of v_mfma_f32_32x32x2_f32 (one k per lane half per instruction; four
// instructions consume the 4 values) and v_mfma_f32_32x32x16_bf16 (8 k per lane
// half) are then a single ds_read_b128 per lane at consecutive addresses.
//
// The writer (global -> LDS) and the reader (LDS -> MFMA) touch the same 16-byte
// slots in different lane patterns. The position of a chunk inside its fragment
// is XOR-swizzled so that every group of 8 lanes (128-byte LDS cycle) and of 16
// lanes (256-byte cycle) lands on distinct 16-byte bank groups for both
// patterns; the host-side test simulates both patterns to prove it.
//
// The A-tile swizzle mixes the chunk index into the low bits (the A writer
// spreads one row's chunks over consecutive lanes); the B-tile swizzle mixes
// row bits 3-4 and (row bit 4 xor block parity) into bits 0-1 and 3 (the B
// writer spreads columns 2 or 4 apart over consecutive lanes). Neither swizzle
// moves a chunk out of its lane half.
#pragma once

#if defined(__HIPCC__)
#define MMSK_HD __host__ __device__ __forceinline__
#else
#define MMSK_HD inline
#endif

namespace mmsk {

// k values per 16-byte chunk, and BF16 terms per FP32 value, for each mode.
MMSK_HD constexpr int mode_lu(int mode) { return mode == 0 ? 4 : 8; }
MMSK_HD constexpr int mode_terms(int mode) { return mode == 0 ? 1 : (mode == 1 ? 2 : 3); }

// log2(8 / chunks-per-row), clamped at 0: how many low row bits the A writer
// places inside one 8-lane group.
MMSK_HD constexpr int a_row_shift(int chunks_per_row) {
    return chunks_per_row >= 8 ? 0 : (chunks_per_row == 4 ? 1 : (chunks_per_row == 2 ? 2 : 3));
}

// Swizzled lane slot (0..63) of chunk c of row r (0..31) of 32-row block rblk.
MMSK_HD constexpr int swizzle_slot(int op, int r, int rblk, int c, int chunks_per_row) {
    const int pos = r + 32 * (c & 1);
    const int s = a_row_shift(chunks_per_row);
    const int mask = op == 0 ? (((c << s) ^ (((r >> s) & 1) << 3)) & 31)
                             : (((r >> 3) & 3) | ((((r >> 4) ^ rblk) & 1) << 3));
    return pos ^ mask;
}

// Byte offset of (row, chunk c, term) inside the operand region of one stage.
// rows = BM (op 0) or BN (op 1); bk = tile depth in k; terms are stored as
// separate fragment arrays, one after another.
MMSK_HD constexpr int chunk_offset(int mode, int op, int rows, int bk, int row, int c, int term) {
    const int lu = mode_lu(mode);
    const int cpr = bk / lu;           // chunks per row
    const int kg_count = cpr / 2;      // 2-chunk k-groups per row
    const int nfrag = (rows / 32) * kg_count;
    const int rblk = row / 32;
    const int r = row % 32;
    const int frag = rblk * kg_count + c / 2;
    return ((term * nfrag + frag) * 64 + swizzle_slot(op, r, rblk, c, cpr)) * 16;
}

// Columns per B staging unit of the tile kernel (a unit is LU k-rows of BW
// consecutive columns). Eight-wave tiles use 2 to keep the staging registers
// within the two-waves-per-SIMD budget. Plain four-wave tiles use 4 (measured
// faster than narrower units). Split four-wave tiles use as few columns as
// still give every thread a unit, so narrow tiles stage B with all threads and
// few registers (measured faster).
MMSK_HD constexpr int b_unit_cols(int mode, int bn, int bk, int waves) {
    const int per_thread = (bk / mode_lu(mode)) * bn / (64 * waves);
    return waves == 8 ? 2 : (mode == 0 || per_thread >= 4 ? 4 : (per_thread >= 2 ? 2 : 1));
}

// Register sets of the tile kernel's global -> LDS staging, i.e. k-blocks in
// flight (measured per mode on one MI355X): two for every plain tile (2-13%
// faster from 512 x 512 x 4096 to 2048^3 and on the narrow shapes), two for
// split6 tiles of up to 64 x 128, one for split3 (the extra registers cost it
// occupancy and 1-6%).
MMSK_HD constexpr int prefetch_sets(int mode, int bm, int bn) {
    return mode == 0 ? 2 : (mode == 2 && bm * bn <= 64 * 128 ? 2 : 1);
}

// Bytes of one operand region (all terms) of one stage.
MMSK_HD constexpr int operand_bytes(int mode, int rows, int bk) {
    return mode_terms(mode) * rows * bk * (mode == 0 ? 4 : 2);
}

// Epilogue staging of the 256 x 256 tile: each wave passes 32 rows of its
// wave tile (wtn = 64 columns) through a row-major block of 32 x epi_pitch(wtn)
// floats. The 4-float pad shifts consecutive rows by 16 bytes, so a store (lane
// l holds row l % 32) and a read (one whole 256-byte row per 16 lanes) each
// touch every 16-byte slot once per cycle at this width.
MMSK_HD constexpr int epi_pitch(int wtn) { return wtn + 4; }
// Float offset in the wave's block of lane `lane`'s 16-byte store of the
// accumulator values 4q..4q+3 of fragment column tj: row lane % 32, columns
// tj * 32 + 8q + 4 * (lane / 32) .. +3.
MMSK_HD constexpr int epi_write_offset(int wtn, int lane, int tj, int q) {
    return (lane & 31) * epi_pitch(wtn) + tj * 32 + 8 * q + 4 * (lane >> 5);
}
// Row and first of the four columns that lane `lane` reads in read step i.
MMSK_HD constexpr int epi_read_row(int wtn, int lane, int i) { return i * (256 / wtn) + lane / (wtn / 4); }
MMSK_HD constexpr int epi_read_col(int wtn, int lane) { return (lane % (wtn / 4)) * 4; }

}  // namespace mmsk
