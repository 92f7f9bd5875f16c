/*
 * Plain-C driver of libclawpack_amd.so's C ABI (include/clawpack_amd.h), with no Python and no
 * PyTorch in the process: device memory from hipMalloc, the control block filled on the host by
 * claw2d_ctrl_init, claw2d_advance enqueued on the default stream, results copied back.  This is
 * what a C, Fortran-via-bind(C) or other-FFI caller of the library does.
 *
 *   drive_advance in.bin out.bin
 *
 * in.bin  : int32 solver, mx, my, order, transverse_waves, limiter, bc, split; float64 dx, dy,
 *           grav, gamma, dt0, tend; then num_eqn planes of my rows of mx doubles (x fastest).
 * out.bin : int32 status, done; int64 numsteps, rejected, attempts; float64 t, dt, cflmax;
 *           then the planes of the final state.
 * Exit code: 0 on success, 1 for a usage or file error, 2 for a library or HIP error (message on
 * stderr).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "../../include/clawpack_amd.h"

#define HIP_OK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                      \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)

#define CLAW_OK(call)                                                                 \
  do {                                                                                \
    int rc_ = (call);                                                                 \
    if (rc_ != 0) {                                                                   \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, claw2d_last_error());       \
      return 2;                                                                       \
    }                                                                                 \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 1;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) { perror(argv[1]); return 1; }
  int32_t hi[8];
  double hd[6];
  if (fread(hi, sizeof(int32_t), 8, f) != 8 || fread(hd, sizeof(double), 6, f) != 6) {
    fprintf(stderr, "short header\n");
    fclose(f);
    return 1;
  }
  claw2d_desc d;
  memset(&d, 0, sizeof(d));
  d.solver = hi[0];
  d.dtype = CLAW_FP64;
  d.mx = hi[1];
  d.my = hi[2];
  d.num_eqn = d.num_waves = (d.solver == CLAW_SOLVER_EULER_4WAVE) ? 4 : 3;
  d.num_ghost = 2;
  d.order = hi[3];
  d.transverse_waves = hi[4];
  for (int p = 0; p < 4; ++p) d.mthlim[p] = hi[5];
  for (int k = 0; k < 2; ++k) {
    d.bc_lower[k] = d.bc_upper[k] = hi[6];
    d.reflect_index[k] = k + 1;
  }
  d.dt_variable = 1;
  d.dx = hd[0];
  d.dy = hd[1];
  d.grav = hd[2];
  d.gamma = hd[3];
  d.cfl_max = 1.0;
  d.cfl_desired = 0.9;
  d.dt_max = 1e99;
  const int split = hi[7];
  const double dt0 = hd[4], tend = hd[5];
  CLAW_OK(claw2d_validate(&d));

  const size_t bytes = claw2d_state_bytes(&d);
  double* q = (double*)malloc(bytes);
  if (!q || fread(q, 1, bytes, f) != bytes) {
    fprintf(stderr, "short state\n");
    fclose(f);
    free(q);
    return 1;
  }
  fclose(f);

  void *q0 = NULL, *q1 = NULL, *ctrl_dev = NULL;
  HIP_OK(hipMalloc(&q0, bytes));
  HIP_OK(hipMalloc(&q1, bytes));
  HIP_OK(hipMalloc(&ctrl_dev, sizeof(claw2d_ctrl)));
  HIP_OK(hipMemcpy(q0, q, bytes, hipMemcpyHostToDevice));
  claw2d_ctrl c;
  CLAW_OK(claw2d_ctrl_init(&d, &c, 0.0, dt0, tend, 1, 0, 100000));
  HIP_OK(hipMemcpy(ctrl_dev, &c, sizeof(c), hipMemcpyHostToDevice));
  /* enqueue attempts in chunks and poll ctrl.done at the chunk ends, as a solver's output loop does */
  for (int chunk = 0; chunk < 1000 && !c.done; ++chunk) {
    if (split)
      CLAW_OK(claw2d_advance_ds(&d, q0, q1, (claw2d_ctrl*)ctrl_dev, 64, NULL));
    else
      CLAW_OK(claw2d_advance(&d, q0, q1, (claw2d_ctrl*)ctrl_dev, 64, NULL));
    HIP_OK(hipMemcpy(&c, ctrl_dev, sizeof(c), hipMemcpyDeviceToHost));
  }
  HIP_OK(hipMemcpy(q, c.cur ? q1 : q0, bytes, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(q0));
  HIP_OK(hipFree(q1));
  HIP_OK(hipFree(ctrl_dev));

  FILE* o = fopen(argv[2], "wb");
  if (!o) { perror(argv[2]); free(q); return 1; }
  const int32_t oi[2] = {c.status, c.done};
  const int64_t ol[3] = {c.numsteps, c.rejected, c.attempts};
  const double od[3] = {c.t, c.dt, c.cflmax};
  int ok = fwrite(oi, sizeof(oi), 1, o) == 1 && fwrite(ol, sizeof(ol), 1, o) == 1 &&
           fwrite(od, sizeof(od), 1, o) == 1 && fwrite(q, 1, bytes, o) == bytes;
  fclose(o);
  free(q);
  if (!ok) { fprintf(stderr, "write failed\n"); return 1; }
  return 0;
}
