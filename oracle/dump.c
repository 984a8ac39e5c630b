/* ORACLE — test infrastructure only. The per-8x8 decision grid of a coded
 * frame (dump_decisions, src/encoder.c:659-704): for every 8x8 cell the
 * block that covers it last in decode order, as 8 bytes: bsize, mode
 * (0 intra, 1 single reference, 2 compound), ref_frame, 0, int16 mv_row,
 * int16 mv_col. Cells no block covers stay zero. */
#include "ocommon.h"

void oracle_decision_grid(const Mav1FrameParams *fp, const Mav1BlockInfo *blocks, const Mav1SbInfo *sbs, uint8_t *grid) {
  const int rows8 = (fp->mi_rows + 1) >> 1, cols8 = (fp->mi_cols + 1) >> 1;
  memset(grid, 0, (size_t)rows8 * cols8 * 8);
  for (int s = 0; s < fp->sb_cols * fp->sb_rows; s++)
    for (int k = 0; k < sbs[s].num_blocks; k++) {
      const Mav1BlockInfo *b = &blocks[sbs[s].first_block + k];
      const int w8 = (Num_4x4_Blocks_Wide[b->bsize] + 1) >> 1, h8 = (Num_4x4_Blocks_High[b->bsize] + 1) >> 1;
      for (int r = b->mi_row >> 1; r < o_min(rows8, (b->mi_row >> 1) + h8); r++)
        for (int c = b->mi_col >> 1; c < o_min(cols8, (b->mi_col >> 1) + w8); c++) {
          uint8_t *cell = grid + ((size_t)r * cols8 + c) * 8;
          const int16_t mr = (int16_t)b->mv_row, mc = (int16_t)b->mv_col;
          cell[0] = (uint8_t)b->bsize;
          cell[1] = !b->is_inter ? 0 : b->ref_frame1 ? 2 : 1;
          cell[2] = (uint8_t)b->ref_frame;
          cell[3] = 0;
          memcpy(cell + 4, &mr, 2);
          memcpy(cell + 6, &mc, 2);
        }
    }
}
