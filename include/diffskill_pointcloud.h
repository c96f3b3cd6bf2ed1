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

#ifdef __cplusplus
}
#endif
#endif
