// arrays.h -- file I/O of the drivers: .npy arrays, and the reference's
// HDF5 files when the drivers are built with libhdf5 (KEVD_HAVE_HDF5).
//
// A file argument is "path.npy" or "path.h5[:/dataset]".  HDF5 datasets are
// written the way the reference writes them: one flat 1-D dataset per array
// (row-major for /U, column-major for /linv, as the callers arrange), FP64,
// FP32 or native unsigned 64-bit integers (ks_arpclm.cu:82-172,
// ks_kevd.cu:134-290).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "npy.h"

bool is_h5(const std::string& spec);
bool have_hdf5();

// Read an array.  HDF5: the dataset of the spec (default `dset`); its
// dims become the shape, floating classes give "<f8" / "<f4", integer
// classes "<i8".  Throws std::runtime_error.
NpyArray load_array(const std::string& spec, const std::string& dset = "/u");

// Write an array.  .npy: shape as given.  HDF5: a flat dataset of
// prod(shape) elements named by the spec (default `dset`); `append` adds
// it to an existing file (creating the file if absent) instead of
// replacing the file.  dtype "<f8", "<f4" or "<u8" ("<i8" is stored as
// unsigned 64-bit, the reference's pivot type).
void save_array(const std::string& spec, const std::string& dtype, const std::vector<int64_t>& shape,
                const void* data, const std::string& dset, bool append = false);

// Points for a driver.  .npy: a 2-D array (rows = points).  HDF5: the
// reference's snapshot file (ks_datagen.py:91-103): /u holds Ntrue * M
// values with the time index fastest (u[g * Ntrue + t]), /t the Ntrue
// times; M = grid_points if > 0, else |u| / |t|.  With delay > 0 the points
// are the ks_arpclm.cu delay windows (point g * N + t is u[g][t .. t+delay),
// N = Ntrue - delay + 1, ks_arpclm.cu:263-268), else the Ntrue snapshots.
// Returns FP64 values (f32 tells whether the file held FP32).
std::vector<double> load_points(const std::string& spec, int64_t delay, int64_t grid_points, int64_t* n,
                                int32_t* dim, bool* f32);

// load_points into a context (FP32 sources through kevd_set_points_f32).
// Returns the library status; file errors print and return KEVD_EINVAL.
struct kevd_ctx;
int set_points_from(kevd_ctx* ctx, const std::string& spec, int64_t delay, int64_t grid_points, int64_t* n,
                    int32_t* dim);
