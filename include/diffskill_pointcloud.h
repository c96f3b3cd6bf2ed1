/*
 * diffskill_pointcloud.h -- C ABI of the point-cloud kernels (gfx950), in the
 * same library as diffskill_mpm.h (libdiffskill_mpm.so).
 *
 * pc_nearest is the nearest-neighbour search under the reference's Chamfer
 * losses: compute_chamfer_loss_kernel (plb/engine/losses/loss.py:183-204,
 * inner loops over all targets) and the chamferdist.ChamferDistance loss
 * built by create_chamfer_loss (plb/engine/taichi_env.py:28-32, a
 * third-party dependency absent here).  The loss values and their adjoints
 * are assembled from its outputs by the host layer.
 *
 * pc_softmin / pc_plan_grad are the pairwise kernels of the Sinkhorn
 * divergence behind the reference's EMD loss, create_emd_loss's
 * geomloss.SamplesLoss(loss='sinkhorn', p=1, blur=0.001)
 * (plb/engine/taichi_env.py:23-26, evaluated at :269-270 and :303; geomloss
 * is a third-party dependency absent here): the soft-min of its dual
 * updates and the transport-plan gradient of its last step.  The annealing
 * loop runs on the host.
 */
#ifndef DISKILL_POINTCLOUD_H
#define DISKILL_POINTCLOUD_H

#ifdef __cplusplus
extern "C" {
#endif

/* For every query point q[i] (nq x 3 floats, device) the nearest target of
 * t (nt x 3 floats, device), by exact fp32 squared Euclidean distance:
 * idx[i] its index (the smallest index among equal distances; -1 when nt is
 * 0) and d2[i] the squared distance (+inf when nt is 0).  work: nq 64-bit
 * device words of scratch, overwritten.  Asynchronous on stream (a
 * hipStream_t); no allocation, no host synchronisation, so the call can be
 * captured into a HIP graph.  Cost: nq * nt distance evaluations.  Returns 0,
 * or -1 with pc_last_error() describing the problem. */
int pc_nearest(const float *q, int nq, const float *t, int nt, unsigned long long *work, int *idx, float *d2, void *stream);
const char *pc_last_error(void);

/* pc_nearest without the workspace argument and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The per-query scratch is a static device array of the
 * library, 2^20 queries (8 MiB) per GPU, held until the process ends; nq
 * above that is refused with an error (-1) before any GPU work.  Every call
 * on one GPU shares that GPU's array (cleared for its nq queries at the
 * start of each call): calls that overlap on one GPU -- on different
 * streams or threads -- race and must be serialised by the caller.  Same
 * outputs as pc_nearest on the same inputs. */
int pc_nearest_noalloc(const float *q, int nq, const float *t, int nt, int *idx, float *d2, void *stream);

/* Bytes of device scratch that pc_softmin and pc_plan_grad need for n rows
 * and m columns (-1 for negative counts). */
long long pc_sinkhorn_work_bytes(int n, int m);

/* out[i] = -eps log sum_j exp(h[j] - C(x_i, y_j) / eps), the weights h
 * given in log2 units (h2 = h * log2(e)), for the n rows x (n x 3
 * floats, device) against the m columns y (m x 3), C = |x - y| (p = 1) or
 * |x - y|^2 / 2 (p = 2); every pair is evaluated (n * m work), the columns
 * split over workgroups whose partial (max, sum) are combined in a fixed
 * order.  work: pc_sinkhorn_work_bytes(n, m) bytes of device scratch.
 * Asynchronous on stream, no allocation.  Returns 0, or -1 with
 * pc_last_error() (m must be positive, eps positive, p 1 or 2). */
int pc_softmin(const float *x, int n, const float *y, int m, const float *h2, float eps, int p, void *work, float *out,
               void *stream);

/* g[i] (n x 3, overwritten) = sum_j exp(h[j] + r[i] - C(x_i, y_j) / eps)
 * dC(x_i, y_j)/dx_i with h = h2 / log2(e), r = r2 / log2(e) (log2 units, as
 * pc_softmin), dC/dx = (x - y) / |x - y| for p = 1 (0 where x = y) and
 * x - y for p = 2: the transport plan's pull on every row point.  Same work
 * buffer, stream and error rules as pc_softmin. */
int pc_plan_grad(const float *x, int n, const float *y, int m, const float *h2, const float *r2, float eps, int p, void *work,
                 float *g, void *stream);

/* Bytes of device scratch pc_sinkhorn_step needs (-1 unless n, m > 0). */
long long pc_sinkhorn_step_work_bytes(int n, int m);

/* One annealing step of the debiased Sinkhorn loop (uniform weights) on the
 * potentials pot = [a_x, b_x, a_y, b_y, pa_x, pb_x, pa_y, pb_y] (device
 * floats, lengths n, n, m, m, n, n, m, m): the four soft-mins
 *   a_y <- softmin_x(log(1/n) + b_x / eps), b_x <- softmin_y(log(1/m) + a_y / eps),
 *   a_x <- softmin_x(log(1/n) + a_x / eps), b_y <- softmin_y(log(1/m) + b_y / eps)
 * all computed from the previous values; mode 0 takes the log weights alone
 * (the first eps) and stores the results, mode 1 stores (previous + result) / 2
 * (the symmetric update of every eps), mode 2 moves the previous values to
 * pa_x.. and stores the results (the last extrapolation).  Two launches on
 * stream, every pair evaluated, fixed combination order, no allocation.
 * Returns 0, or -1 with pc_last_error(). */
int pc_sinkhorn_step(const float *x, int n, const float *y, int m, float *pot, float eps, int p, int mode, void *work,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif
