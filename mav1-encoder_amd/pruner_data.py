"""Training samples for the learned pruner from the plain path's decisions.

One sample per 64x64 superblock of a coded frame (superblocks lying wholly
inside the visible frame):
  - input: the source luma / 255 and the residual (source minus the
    co-located luma of the frame's LAST reference picture, zero on key
    frames) / 255, 64 x 64 x 2 floats; q = base_q_idx / 255;
  - partition class (the superblock's decision, from its blocks): 0 one
    64x64 block; 1 the smallest block is 32 or 64 samples on its short side
    (64x32, 32x64, 32x32 splits); 2 the smallest is 16 on its short side;
    3 blocks below 16 on a side;
  - mode class: the class of the inter / intra decision covering most of the
    superblock's area: 2 * kind + (motion vector non-zero), kind 0 intra
    (mv bit 0), 1 LAST, 2 the second reference, 3 compound. (The reference's
    analysis evaluates NEWMV per reference and compound only,
    src/kernels/k_analysis.h:256-311, so this is the shortlist the search
    can be pruned to.)
The decisions come from the coded frames the sequence driver returns
(Mav1HipCodedFrame: blocks, superblock records, final pictures).
"""
import numpy as np

import synth

# Block widths / heights in samples by bsize (BLOCK_4X4 .. BLOCK_64X16).
BW = [4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 4, 16, 8, 32, 16, 64]
BH = [4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 64, 32, 64, 128, 64, 128, 16, 4, 32, 8, 64, 16]
BLOCK_64X64 = 12


def partition_class(blocks):
    """blocks: the superblock's block records (n x 27 int32)."""
    bs = blocks[:, 2]
    if len(bs) == 1 and bs[0] == BLOCK_64X64:
        return 0
    short = min(min(BW[b], BH[b]) for b in bs)
    return 1 if short >= 32 else 2 if short >= 16 else 3


def mode_class(blocks):
    area = np.zeros(8)
    for b in blocks:
        a = BW[b[2]] * BH[b[2]]
        if not b[3]:
            area[0] += a
            continue
        kind = 3 if b[24] else 1 if b[23] == 1 else 2  # ref_frame1 (compound), ref_frame LAST_FRAME = 1
        area[2 * kind + int(b[8] != 0 or b[9] != 0)] += a
    return int(np.argmax(area))


def samples_from_stream(cfg, planes, coded):
    """planes: the stream's I420 frames in display order; coded: its coded
    frames in coding order (dicts of mav1hip.Sequence.receive). Returns
    (x [N, 64, 64, 2] float32, q [N] float32, labels [N, 2] int32)."""
    geo = synth.Geometry(cfg.width, cfg.height, cfg.tile_cols_log2)
    slots = [None] * 8
    xs, qs, ls = [], [], []
    for c in coded:
        fp = c["fp"]
        src = synth.pad_frame(geo, planes[c["disp"]])
        Y = src[:geo.stride[0] * geo.plane_h[0]].reshape(geo.plane_h[0], geo.stride[0]).astype(np.float32)
        ref = None
        if not c["is_key"] and c["last_slot"] >= 0 and slots[c["last_slot"]] is not None:
            ref = slots[c["last_slot"]][:geo.stride[0] * geo.plane_h[0]].reshape(geo.plane_h[0], geo.stride[0])
            ref = ref.astype(np.float32)
        for s in range(fp.sb_cols * fp.sb_rows):
            r, col = divmod(s, fp.sb_cols)
            if (r + 1) * 64 > cfg.height or (col + 1) * 64 > cfg.width:
                continue
            first, n = int(c["sbs"][s, 0]), int(c["sbs"][s, 1])
            blk = c["blocks"][first:first + n]
            ys, xs0 = slice(r * 64, r * 64 + 64), slice(col * 64, col * 64 + 64)
            luma = Y[ys, xs0]
            res = luma - ref[ys, xs0] if ref is not None else np.zeros_like(luma)
            xs.append(np.stack([luma / 255.0, res / 255.0], axis=-1))
            qs.append(fp.base_q_idx / 255.0)
            ls.append((partition_class(blk), mode_class(blk)))
        for k in range(8):
            if (c["refresh_frame_flags"] >> k) & 1:
                slots[k] = c["recon"]
    return (np.asarray(xs, np.float32).reshape(-1, 64, 64, 2), np.asarray(qs, np.float32),
            np.asarray(ls, np.int32).reshape(-1, 2))
