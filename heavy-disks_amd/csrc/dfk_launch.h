// Host-callable launchers of the HIP kernels in dfk_kernels.hip.
#ifndef DFK_LAUNCH_H
#define DFK_LAUNCH_H

#include <hip/hip_runtime.h>

#include "dfk_internal.h"
#include "../../include/diskfft.h"

hipError_t dfk_launch_control(hipStream_t s, int ndisk, const DiskDev *disks, DiskCtl *ctl,
                              const double *pcfl, const double *pcm, DfkSampleDev *samples,
                              int cap, int par, DiskCtl *mirror = nullptr);
hipError_t dfk_launch_diag(hipStream_t s, int nblk, const DiskDev *disks, const int2 *map,
                           int base, const DiskCtl *ctl, const double *st, long long ntot,
                           const double *geom, const double *trig, double *pcfl, double *pcm);
// nctl > 0 appends one control workgroup per disk ctl0 .. ctl0+nctl-1 to the launch.
// mirror (here and in the control launches): null, or a host-visible copy of the
// control blocks that the control workgroups also write (the last iteration of a
// captured run).
// inplace: 0 the cooperative row transforms, 1 the in-place ones (k_rows_*_group) with
// one complex row per workgroup (pr = 1 for every disk), 2 in place with batches,
// 3 the register two-pass kernels (k_rows_*_reg; every Nphi <= 512, 64 threads,
// dfk_rows_reg_pairs row pairs per workgroup), 4 the same with Nphi up to 1024,
// 5 the register three-pass kernels (k_rows_*_reg3, one row pair per workgroup).
hipError_t dfk_launch_r2c(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                          const int2 *map, int base, DiskCtl *ctl, const double *st,
                          long long ntot, const double *geom, const double2 *tw, double2 *spec,
                          int par, int nctl, int ctl0, const double *pcfl, const double *pcm,
                          DfkSampleDev *samples, int cap, int inplace, DiskCtl *mirror = nullptr);
// Loop control riding in a column-convolution launch: nctl trailing
// workgroups run the control of disks ctl0 .. ctl0 + nctl - 1 (nctl = 0: none).
struct DfkConvCtl {
  int nctl = 0, ctl0 = 0;
  const double *pcfl = nullptr, *pcm = nullptr;
  DfkSampleDev *samples = nullptr;
  int cap = 0, par = 0;
  DiskCtl *mirror = nullptr;
};
hipError_t dfk_launch_conv(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                           const int2 *map, int base, DiskCtl *ctl, const double *khat,
                           const double2 *tw, double2 *spec, const DfkConvCtl &c);
hipError_t dfk_launch_conv_wave(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                                const int2 *map, int base, DiskCtl *ctl, const double *khat,
                                const double2 *tw, double2 *spec, const DfkConvCtl &c);
// batched: some disk of the launch has cb > 1 (else the one-column build).
hipError_t dfk_launch_conv_group(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                                 const int2 *map, int base, DiskCtl *ctl, const double *khat,
                                 const double2 *tw, double2 *spec, int batched, const DfkConvCtl &c);
// Two-pass register column convolution (transform lengths M = 128, 192, 256,
// 384, 512, 768, 1024): columns per 64-thread workgroup (0: M unsupported),
// dynamic LDS bytes a disk of that M needs, and the launch (lds: the largest
// need among the launch's disks).
int dfk_conv_reg_cols(int M);
size_t dfk_conv_reg_lds(int M);
// Three-pass register column convolution (column mode 4, one column per
// 64-thread workgroup in the same launch): whether M (= 256, 384, 512, 768,
// 1024) has it, and the dynamic LDS bytes a disk of that M needs.
int dfk_conv_reg3_ok(int M);
size_t dfk_conv_reg3_lds(int M);
// Three-pass register row transforms (row launch code 5: one row pair per
// 64-thread workgroup): whether a row of n points (128, 192, 256, 384, 512)
// has them, and their dynamic LDS bytes.
int dfk_rows_reg3_ok(int n);
size_t dfk_rows_reg3_lds();
// Register row transforms: row pairs per 64-thread workgroup for a row of n
// points (0: n unsupported) and the dynamic LDS bytes of such a disk.
int dfk_rows_reg_pairs(int n);
size_t dfk_rows_reg_lds(int n);
hipError_t dfk_launch_conv_reg(hipStream_t s, int nblk, size_t lds, const DiskDev *disks,
                               const int2 *map, int base, DiskCtl *ctl, const double *khat,
                               const double2 *tw, double2 *spec, const DfkConvCtl &c);
// Fused row kernel (row c2r + update + row r2c) over row blocks of ti x Nphi,
// nthr = 256, 512 or 1024 threads; V-hat read from vin, the new state's
// S-hat written to sout.
hipError_t dfk_launch_rows_fused(hipStream_t s, int nblk, int nthr, size_t lds, int g2,
                                 const DiskDev *disks, const int2 *map, int base, const DiskCtl *ctl,
                                 double *st, long long ntot, const double2 *vin, double2 *sout,
                                 const double *geom, const double2 *tw, const double *trig,
                                 double *pcfl, double *pcm, int par);
// Its dynamic LDS bytes for a disk, and the largest Nphi it transforms with
// nthr threads.
size_t dfk_fused_lds(int ti, int nphi, int tw_lds);
int dfk_fused_max_nphi(int nthr);
hipError_t dfk_launch_c2r(hipStream_t s, int nblk, int nthr, size_t lds, const DiskDev *disks,
                          const int2 *map, int base, const DiskCtl *ctl, const double2 *spec,
                          const double *geom, const double2 *tw, double *phi, int inplace);
// g2: every disk of the launch has gamma = 2 (EOS without pow).
hipError_t dfk_launch_update(hipStream_t s, int nblk, int nthr, size_t lds, int garr, int g2,
                             const DiskDev *disks,
                             const int2 *map, int base, const DiskCtl *ctl, double *st,
                             long long ntot, const double *phi, const double *gR,
                             const double *gp, int has_gR, int has_gp, const double *geom,
                             const double *trig, double *pcfl, double *pcm, int par);
hipError_t dfk_launch_update_stream(hipStream_t s, int nblk, int garr, int g2, const DiskDev *disks,
                                    const int2 *map, int base, const DiskCtl *ctl, double *st,
                                    long long ntot, const double *phi, const double *gR,
                                    const double *gp, int has_gR, int has_gp, const double *geom,
                                    const double *trig, double *pcfl, double *pcm, int par);
hipError_t dfk_launch_khat(hipStream_t s, const DiskDev *D, int nrow_blk, int ncol_blk,
                           size_t lds_rows, size_t lds_cols, const double *ktab,
                           const double2 *tw, double2 *tmp, double *khat);
hipError_t dfk_launch_forces(hipStream_t s, const DiskDev *D, long long N, const double *phi,
                             const double *geom, double *gR, double *gp);
hipError_t dfk_launch_stage(hipStream_t s, const DiskDev *D, long long N, int stage, double *st,
                            long long ntot, int cur, double dt, const double *gR,
                            const double *gp, const double *geom, double *F);

// Scratch of the no-allocation entry point (dfk_noalloc.hip).  init: once
// per device, with that device current (dfk_create); returns 0, or 1 with a
// message in err.  ctl: the device's zeroed control block (null before
// init); spectrum: its half-spectrum area of DFK_NOALLOC_MAX_SPECTRUM double2.
int dfk_noalloc_init(int device, char *err, size_t errlen);
void *dfk_noalloc_ctl(int device);
void *dfk_noalloc_spectrum(int device);

#endif
