/* hipsp: products of a beamlet dose matrix with many weight vectors on one MI355X.
 *
 * Replaces the planning loop's DijResult.dose_for_weights
 * (pyradmc/scoring/dij.py:251-260), which forms one weight vector's dose on the host,
 * with a device product for a whole set of vectors: D[voxel][k] = sum over beamlets j
 * of Dij[voxel][j] * W[j][k]. The CSC matrix (rows = scoring voxels, the layout
 * run_dij returns) is uploaded once and kept on the device, every row in ascending
 * column order. In HSP_F64 each output is the float64 sum the reference forms
 * (float64 product of the entry and the weight, rounded, then added in ascending
 * column order; no fused multiply-add), bit for bit.
 */
#ifndef HIPSP_H
#define HIPSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSP_OK 0
#define HSP_EINVAL -1
#define HSP_EHIP -2
#define HSP_ENOMEM -5

/* Element types: of the stored values (hsp_create) and of a product's arithmetic. */
#define HSP_F64 0 /* float64; a product is then the reference's sum bit for bit */
#define HSP_F32 1 /* float32; a product uses fused multiply-add */

typedef struct hsp_matrix hsp_matrix;

/* Upload a CSC matrix (indptr [n_cols + 1], row indices [nnz] in [0, n_rows), values
 * [nnz] of value_type: HSP_F64 for double, as the reference's Dij holds them, or
 * HSP_F32 for float; host arrays) to `device` and build its row-major forms there
 * (the matrix's counterpart: DijResult's indptr / indices / dose,
 * pyradmc/scoring/dij.py:194-219). Entries equal to zero contribute nothing to a
 * product; repeated (row, column) entries add in their CSC order. */
int hsp_create(int device, long long n_rows, long long n_cols, const int64_t *indptr,
               const int64_t *indices, const void *values, int value_type, hsp_matrix **out);
void hsp_destroy(hsp_matrix *m);
const char *hsp_last_error(hsp_matrix *m);

/* out[r - row_begin][k] = sum_j D[r][j] W[j][k] for rows [row_begin, row_end).
 * weights: host float64 [n_cols][n_w] (rounded to float32 in HSP_F32).
 * out_device: device memory of (row_end - row_begin) * n_w elements of the
 * precision's type, or NULL for a buffer the matrix keeps (read it with hsp_fetch).
 * kernel_ms (optional) receives the product kernel's time. Blocking. */
int hsp_apply(hsp_matrix *m, const double *weights, int n_w, long long row_begin, long long row_end,
              int precision, void *out_device, double *kernel_ms);

/* As hsp_apply with the weights already on the device (float64 for HSP_F64,
 * float32 for HSP_F32, [n_cols][n_w]). */
int hsp_apply_device(hsp_matrix *m, const void *weights_device, int n_w, long long row_begin,
                     long long row_end, int precision, void *out_device, double *kernel_ms);

/* Rows [row_begin, row_end) of the last product kept in the internal buffer, to host
 * memory [rows][n_w] of that product's element type. */
int hsp_fetch(hsp_matrix *m, long long row_begin, long long row_end, void *host);

/* The device CSR, copied to host arrays (row_ptr [n_rows + 1], col, val [nnz] of the
 * matrix's value type). */
int hsp_csr(hsp_matrix *m, int64_t *row_ptr, int32_t *col, void *val);

/* Size of the row-block form: the number of (block, column) entries over all blocks
 * of 8 rows. */
long long hsp_block_entries(hsp_matrix *m);

/* out[j] = sum_i D[i][j] v[i] for every column j: the transpose product the planning
 * objective's gradient needs (pyradmc/study.py:93-114, D_roi.T @ residual). v: host
 * float64 [n_rows]; out: host float64 [n_cols]. Each column is summed by one
 * workgroup in a fixed order (strided partial sums, then a tree), so the result is
 * reproducible but not the sequential sum bit for bit. kernel_ms optional. Blocking. */
int hsp_apply_transpose(hsp_matrix *m, const double *v, double *out, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
