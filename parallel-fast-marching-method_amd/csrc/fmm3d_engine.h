// fmm3d_engine.h -- continuous batching inside one context (library
// internal; the serving loop in fmm3d_server.hip drives it).
//
// A context opened as an engine holds `cap` slots of one precision.  Slots
// are admitted one source at a time while the column walk keeps sweeping:
// a new slot starts at the current sweep number with every patch column
// active; a slot converges when a sweep changes none of its values; its
// receiver times are read and the slot is retired and free again.  Slots
// never interact, so a source gives the same field whichever others share
// the sweeps (up to float32 rounding of the concurrent halo reads).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fmm3d.h"

#define FMM3D_HIDDEN __attribute__((visibility("hidden")))

FMM3D_HIDDEN int fmm3d_engine_open(fmm3d_ctx* c, int cap, int precision, int order);
// one source into a free slot: its fixed set (node index, time), whether
// its column walk alternates, and a point-source position (or null)
FMM3D_HIDDEN int fmm3d_engine_admit(fmm3d_ctx* c, int slot, const int64_t* node, const double* tv, int64_t n,
                                    int walkalt, const double* src3);
// a teleseismic source into a free slot from its bottom face (nnx*nnz
// float64, phi fastest, values > 0 fixed); queued behind the previous
// launch, no wait; `face` must stay valid until the next launch completes
FMM3D_HIDDEN int fmm3d_engine_admit_face(fmm3d_ctx* c, int slot, const double* face);
// queue `nsweeps` sweeps over the admitted slots on the context's stream
FMM3D_HIDDEN int fmm3d_engine_launch(fmm3d_ctx* c, int nsweeps);
// wait for the queued sweeps; slots that converged go to `done`, slots that
// hit the sweep cap go to `failed`
FMM3D_HIDDEN int fmm3d_engine_collect(fmm3d_ctx* c, std::vector<int>& done, std::vector<int>& failed);
FMM3D_HIDDEN int fmm3d_engine_receivers(fmm3d_ctx* c, int slot, int nrc, const double* rcv, float* out);
// receiver times of n slots in one launch and one wait: slot slots[k] at
// its nrc[k] receivers rcv[k] into out[k]
FMM3D_HIDDEN int fmm3d_engine_receivers_many(fmm3d_ctx* c, int n, const int* slots, const int* nrc,
                                             const double* const* rcv, float* const* out);
// the rays and Frechet rows of n slots in one pass (slot slots[k] at its
// nrc[k] receivers rcv[k]): rows slot after slot, rp gets sum(nrc)+1 offsets
FMM3D_HIDDEN int fmm3d_engine_frechet_many(fmm3d_ctx* c, int n, const int* slots, const int* nrc,
                                           const double* const* rcv, std::vector<int64_t>& rp,
                                           std::vector<int32_t>& cols, std::vector<float>& vals);
FMM3D_HIDDEN int fmm3d_engine_retire(fmm3d_ctx* c, int slot);
FMM3D_HIDDEN int fmm3d_engine_active(const fmm3d_ctx* c);
// A host-only view of a grid and vertex model (no device memory, no
// stream): what the refined neighbourhood of a point source needs, for a
// thread that prepares fixed sets beside the one that sweeps.
FMM3D_HIDDEN fmm3d_ctx* fmm3d_engine_shadow(const fmm3d_grid_desc* g, const double* velv, int device);
FMM3D_HIDDEN void fmm3d_engine_shadow_free(fmm3d_ctx* shadow);
// fixed sets of n point sources (src: n records of height, colatitude,
// longitude), their refined neighbourhoods solved in float64 in groups on
// *fine, a refined-grid context created or resized here and destroyed by
// the caller with fmm3d_destroy; node[q] / tv[q] receive source q's set.
// Reads only the grid, model and device of c
FMM3D_HIDDEN int fmm3d_engine_point_sets(const fmm3d_ctx* c, fmm3d_ctx** fine, std::string& err, int n,
                                         const double* src, int refine, int radius, int order,
                                         std::vector<std::vector<int64_t>>& node,
                                         std::vector<std::vector<double>>& tv);
// device bytes: the model and metric tables; one slot of a precision
FMM3D_HIDDEN size_t fmm3d_engine_model_bytes(const fmm3d_grid_desc* g);
FMM3D_HIDDEN size_t fmm3d_engine_slot_bytes(const fmm3d_grid_desc* g, int precision);
