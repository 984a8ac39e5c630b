/* ORACLE — test infrastructure only. Transform / quantiser functions
 * shared by the oracle's reconstruction. */
#ifndef MAV1_ORACLE_TX_H
#define MAV1_ORACLE_TX_H
#include <stdint.h>

int o_row_kind(int t);
int o_col_kind(int t);
int o_flip_ud(int t);
int o_flip_lr(int t);
void o_itx_1d(int32_t *T, int kind, int n, int r);
void o_itx_2d(const int32_t *dq, int32_t *res, int txSz, int txType, int lossless);
void o_fwd_2d(const int32_t *resid, int32_t *out, int txSz, int txType, const int32_t *mats);
int o_dq_denom(int txSz);
void o_quantize(const int32_t *coef, int32_t *quant, int txSz, int dcQ, int acQ);
void o_dequant(const int32_t *quant, int32_t *dq, int txSz, int dcQ, int acQ);
void o_fwht4x4(const int32_t *resid, int32_t *out);
#endif
