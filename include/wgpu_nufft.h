/*
 * Host-array C interface of the MI355X NUFFT engine, symbol- and
 * layout-compatible with the reference's C ABI (capi/include/wgpu_nufft.h),
 * so that the reference's Fortran, MATLAB and Julia wrappers can link
 * against libmnufft.so unchanged.
 *
 * Each declaration below replaces the reference declaration at the cited
 * line of capi/include/wgpu_nufft.h.  Behaviour that differs on MI355X:
 *  - GPU plans run on the HIP device 0 of the process (the adapter fields of
 *    the options must be unset or match it);
 *  - double-precision plans run native f64 (PRECISION_DF64 is accepted and
 *    runs f64);
 *  - there is no CPU engine: BACKEND_CPU fails with ERROR_GPU_UNAVAILABLE.
 * Arrays are host memory owned by the caller; complex values interleaved;
 * modes axis-0 fastest; batches transform-major.
 */
#ifndef WGPU_NUFFT_H
#define WGPU_NUFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (wgpu_nufft.h:41-52). */
#define WGPU_NUFFT_SUCCESS 0
#define WGPU_NUFFT_ERROR_INVALID_ARGUMENT 1
#define WGPU_NUFFT_ERROR_PLAN 2
#define WGPU_NUFFT_ERROR_GPU_UNAVAILABLE 3
#define WGPU_NUFFT_ERROR_GPU 4
#define WGPU_NUFFT_ERROR_INTERNAL 5

/* Backends (wgpu_nufft.h:54-57). */
#define WGPU_NUFFT_BACKEND_AUTO 0
#define WGPU_NUFFT_BACKEND_GPU 1
#define WGPU_NUFFT_BACKEND_CPU 2

/* Arithmetic (wgpu_nufft.h:59-66). */
#define WGPU_NUFFT_PRECISION_AUTO 0
#define WGPU_NUFFT_PRECISION_F64 1
#define WGPU_NUFFT_PRECISION_DF64 2
#define WGPU_NUFFT_PRECISION_F32 3

/* Mode orders (wgpu_nufft.h:68-70). */
#define WGPU_NUFFT_MODE_ORDER_CENTERED 0
#define WGPU_NUFFT_MODE_ORDER_FFT 1

#define WGPU_NUFFT_ADAPTER_NAME_SIZE 256
#define WGPU_NUFFT_PCI_BUS_ID_SIZE 32

/* Options, 320 bytes (wgpu_nufft.h:76-87). */
typedef struct wgpu_nufft_opts {
    int32_t backend;
    int32_t precision;
    int32_t mode_order;
    int32_t threads;
    double sigma;
    int32_t adapter_index;
    char adapter_name[WGPU_NUFFT_ADAPTER_NAME_SIZE];
    char adapter_pci_bus_id[WGPU_NUFFT_PCI_BUS_ID_SIZE];
} wgpu_nufft_opts;

/* Adapter description (wgpu_nufft.h:118-126). */
typedef struct wgpu_nufft_adapter {
    char name[WGPU_NUFFT_ADAPTER_NAME_SIZE];
    char backend[16];
    char device_type[16];
    char pci_bus_id[WGPU_NUFFT_PCI_BUS_ID_SIZE];
    int32_t is_default;
} wgpu_nufft_adapter;

/* wgpu_nufft.h:91 */
void wgpu_nufft_default_opts(wgpu_nufft_opts *opts);
/* wgpu_nufft.h:135; lists the HIP devices (backend "HIP"). */
int wgpu_nufft_list_adapters(wgpu_nufft_adapter *adapters, int32_t capacity, int32_t *count);
/* wgpu_nufft.h:138 */
const char *wgpu_nufft_last_error(void);
/* wgpu_nufft.h:141 */
const char *wgpu_nufft_version(void);
/* wgpu_nufft.h:145; "<device name> (HIP)". */
const char *wgpu_nufft_gpu_name(void);
/* wgpu_nufft.h:149; releases the one-call plan cache. */
void wgpu_nufft_shutdown(void);

typedef struct wgpu_nufft_plan_s *wgpu_nufft_plan;

/* wgpu_nufft.h:162-164 */
int wgpu_nufft_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                        int64_t ntrans, double eps, wgpu_nufft_plan *plan,
                        const wgpu_nufft_opts *opts);
/* wgpu_nufft.h:172-174 */
int wgpu_nufft_setpts(wgpu_nufft_plan plan, int64_t M, const double *x, const double *y,
                      const double *z, int64_t N, const double *s, const double *t,
                      const double *u);
/* wgpu_nufft.h:180-181 */
int wgpu_nufft_setpts_nd(wgpu_nufft_plan plan, int64_t M, const double *points, int64_t N,
                         const double *targets);
/* wgpu_nufft.h:188 */
int wgpu_nufft_execute(wgpu_nufft_plan plan, double *c, double *f);
/* wgpu_nufft.h:191 */
void wgpu_nufft_destroy(wgpu_nufft_plan plan);
/* wgpu_nufft.h:195-196 */
int wgpu_nufft_plan_backend(wgpu_nufft_plan plan);
int wgpu_nufft_plan_precision(wgpu_nufft_plan plan);

/* One-call transforms (wgpu_nufft.h:204-226). */
int wgpu_nufft1d1(int64_t M, const double *x, const double *c, int32_t isign, double eps,
                  int64_t ms, double *f, const wgpu_nufft_opts *opts);
int wgpu_nufft1d2(int64_t M, const double *x, double *c, int32_t isign, double eps,
                  int64_t ms, const double *f, const wgpu_nufft_opts *opts);
int wgpu_nufft1d3(int64_t M, const double *x, const double *c, int32_t isign, double eps,
                  int64_t N, const double *s, double *f, const wgpu_nufft_opts *opts);
int wgpu_nufft2d1(int64_t M, const double *x, const double *y, const double *c, int32_t isign,
                  double eps, int64_t ms, int64_t mt, double *f, const wgpu_nufft_opts *opts);
int wgpu_nufft2d2(int64_t M, const double *x, const double *y, double *c, int32_t isign,
                  double eps, int64_t ms, int64_t mt, const double *f,
                  const wgpu_nufft_opts *opts);
int wgpu_nufft2d3(int64_t M, const double *x, const double *y, const double *c, int32_t isign,
                  double eps, int64_t N, const double *s, const double *t, double *f,
                  const wgpu_nufft_opts *opts);
int wgpu_nufft3d1(int64_t M, const double *x, const double *y, const double *z, const double *c,
                  int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu, double *f,
                  const wgpu_nufft_opts *opts);
int wgpu_nufft3d2(int64_t M, const double *x, const double *y, const double *z, double *c,
                  int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu,
                  const double *f, const wgpu_nufft_opts *opts);
int wgpu_nufft3d3(int64_t M, const double *x, const double *y, const double *z, const double *c,
                  int32_t isign, double eps, int64_t N, const double *s, const double *t,
                  const double *u, double *f, const wgpu_nufft_opts *opts);

typedef struct wgpu_nufftf_plan_s *wgpu_nufftf_plan;

/* Single precision (wgpu_nufft.h:232-269). */
int wgpu_nufftf_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                         int64_t ntrans, double eps, wgpu_nufftf_plan *plan,
                         const wgpu_nufft_opts *opts);
int wgpu_nufftf_setpts(wgpu_nufftf_plan plan, int64_t M, const float *x, const float *y,
                       const float *z, int64_t N, const float *s, const float *t,
                       const float *u);
int wgpu_nufftf_setpts_nd(wgpu_nufftf_plan plan, int64_t M, const float *points, int64_t N,
                          const float *targets);
int wgpu_nufftf_execute(wgpu_nufftf_plan plan, float *c, float *f);
void wgpu_nufftf_destroy(wgpu_nufftf_plan plan);
int wgpu_nufftf_plan_backend(wgpu_nufftf_plan plan);
int wgpu_nufftf_plan_precision(wgpu_nufftf_plan plan);

int wgpu_nufftf1d1(int64_t M, const float *x, const float *c, int32_t isign, double eps,
                   int64_t ms, float *f, const wgpu_nufft_opts *opts);
int wgpu_nufftf1d2(int64_t M, const float *x, float *c, int32_t isign, double eps, int64_t ms,
                   const float *f, const wgpu_nufft_opts *opts);
int wgpu_nufftf1d3(int64_t M, const float *x, const float *c, int32_t isign, double eps,
                   int64_t N, const float *s, float *f, const wgpu_nufft_opts *opts);
int wgpu_nufftf2d1(int64_t M, const float *x, const float *y, const float *c, int32_t isign,
                   double eps, int64_t ms, int64_t mt, float *f, const wgpu_nufft_opts *opts);
int wgpu_nufftf2d2(int64_t M, const float *x, const float *y, float *c, int32_t isign,
                   double eps, int64_t ms, int64_t mt, const float *f,
                   const wgpu_nufft_opts *opts);
int wgpu_nufftf2d3(int64_t M, const float *x, const float *y, const float *c, int32_t isign,
                   double eps, int64_t N, const float *s, const float *t, float *f,
                   const wgpu_nufft_opts *opts);
int wgpu_nufftf3d1(int64_t M, const float *x, const float *y, const float *z, const float *c,
                   int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu, float *f,
                   const wgpu_nufft_opts *opts);
int wgpu_nufftf3d2(int64_t M, const float *x, const float *y, const float *z, float *c,
                   int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu,
                   const float *f, const wgpu_nufft_opts *opts);
int wgpu_nufftf3d3(int64_t M, const float *x, const float *y, const float *z, const float *c,
                   int32_t isign, double eps, int64_t N, const float *s, const float *t,
                   const float *u, float *f, const wgpu_nufft_opts *opts);

#ifdef __cplusplus
}
#endif

#endif /* WGPU_NUFFT_H */
