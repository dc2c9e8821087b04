/* chemps2_amd MPS checkpoints: the HDF5 layout of CheMPS2's DMRG::saveMPS / loadDIM /
 * loadMPS (DMRGmpsio.cpp:27-162), in libchemps2_amd_io.so (links libhdf5 and
 * libchemps2_amd). Return codes as in chemps2_amd.h. */
#ifndef CHEMPS2_AMD_IO_H
#define CHEMPS2_AMD_IO_H

#include <stdint.h>

#include "chemps2_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DMRG::saveMPS (DMRGmpsio.cpp:27-94): /Convergence/Converged_yn, every
 * /VirtDim_<b>_<N>_<2S>_<I>/Value over SyBookkeeper's sector ranges, /MPS_<site>/Values. */
int chm_mps_save(chm_ctx* ctx, const char* path, int converged);

/* DMRG::loadDIM (DMRGmpsio.cpp:96-126): the virtual dimensions of the file become the
 * current sectors of every boundary. */
int chm_mps_load_dims(chm_ctx* ctx, const char* path);

/* DMRG::loadMPS (DMRGmpsio.cpp:128-162): every site tensor (sizes must match the
 * current sectors) uploaded; *converged (may be null) from /Convergence. */
int chm_mps_load(chm_ctx* ctx, const char* path, int* converged);

/* Host-only forms (no context, no GPU): dims[q] for the q-th (b, N, 2S, I) of
 * SyBookkeeper's ranges in the reference's loop order (DMRGmpsio.cpp:47-52). */
int chm_io_write(const char* path, int L, int N, int two_s, int nirrep, int converged, const int* dims,
                 const double* const* mps, const int64_t* mps_size);
int chm_io_read_dims(const char* path, int L, int N, int two_s, int nirrep, int* dims, int* n_dims);
int chm_io_read_site(const char* path, int site, double* values, int64_t size, int* converged);

#ifdef __cplusplus
}
#endif

#endif
