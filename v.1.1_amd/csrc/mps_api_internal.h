// mps_api_internal.h -- host-side preparation shared by the single and batched forms.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mps_gpu.h"

namespace mpsh {

struct HostPrep {
    int32_t N = 0, Ne = 0, nen = 0, law = 0;
    std::vector<double> X;
    std::vector<int32_t> conn, rowptr, col;  // conn and the per-element arrays in stored order
    std::vector<int32_t> eorig;              // input id of each stored element (empty: input order)
    // patch tables (see mps_common.h)
    int32_t n_patch = 0, n_part = 0;
    std::vector<int32_t> poff, nrow, ncol;
    std::vector<uint16_t> prow, pcol;
    std::vector<int32_t> pnd;  // node id of each partial (the patches' distinct nodes, ascending per patch)
    bool per_elem = false;
    std::vector<double> mat, cel, rho;  // per element
    double pa = 0, pb = 0, c0 = 0, rho0 = 0;  // uniform
    std::vector<uint32_t> ninfo;
    std::vector<double> bcval;
    int32_t bc_mode = 0, vel_ramp = 0;
    mps_solver_cfg cfg{};
};

std::string fmt(const char* f, ...);
void material_consts(int law, double E, double nu, double rho, double* pa, double* pb, double* c);
// Validates the inputs and builds CSR, Dirichlet records and material constants.
int prepare(const mps_mesh* mesh, const mps_material* mat, const mps_bcs* bcs, const mps_solver_cfg* cfg,
            HostPrep* hp, std::string* err);

}  // namespace mpsh

// Device bytes of one solve in a batch arena (mps_batch.hip).
size_t mps_instance_bytes(const mpsh::HostPrep& hp, int precision);
// Batch internals used by the server (mps_batch.hip).
int mps_batch_check(mps_batch b, const mpsh::HostPrep& hp, std::string* why);
size_t mps_batch_footprint(mps_batch b, const mpsh::HostPrep& hp);
void mps_batch_room(mps_batch b, int* free_slots, size_t* largest);
int mps_batch_add_prepared(mps_batch b, const mpsh::HostPrep& hp, int32_t* slot_out);
// Launch bookkeeping and tuning used by the server's online tuner (mps_batch.hip).
double mps_batch_launch_work(mps_batch b, int64_t launches);
int64_t mps_batch_active_elements(mps_batch b);
double mps_batch_last_launch_ms(mps_batch b);
void mps_batch_tuning(mps_batch b, int* elem_block, int* grid_mult);
int mps_batch_set_tuning(mps_batch b, int elem_block, int grid_mult);
