// Shared definitions of the device legalizer (host and device side).
#pragma once

#include <cstdint>

namespace lgdev {

// Device-resident iteration control.  Written only by single-thread kernels
// (k_scan, k_finish, setup kernels) and by atomic accumulations of the
// partition kernel; every other kernel only reads it.
struct Ctl {
  int32_t iter;          // index of the iteration being executed
  int32_t done;          // 1: converged or max_admm_iter reached -> kernels exit
  int32_t converged;
  int32_t conv_iter;     // iteration at which overflow reached the threshold
  int32_t npad;          // pow2(max cells in a valid partition) for this iteration
  int32_t max_cnt;
  int32_t exceptional;   // next iteration runs the adaptive search
  int32_t last_slow;
  int32_t min_ovf;
  int32_t error;         // nonzero: a device-side limit was exceeded (see kErr*)
  int32_t adaptive_calls;
  int32_t wrong_row;     // cells on a row of the wrong rail parity at solve start
  int32_t n_work[2];     // partitions listed in Args::work per sort-size class
  int32_t ovf_window_n;  // entries in the stagnation window
  int32_t ovf_window[64];  // kMaxStagnationWindow
  float rho;
  float lambda0;
  float step;
  int32_t n_huge;        // partitions above kMaxNpadLds slots sorting in Args::huge this step
  unsigned long long ovf_acc;     // this iteration's overflow (sum over partitions)
  unsigned long long last_ovf;
  unsigned long long demand_sum;  // exact integer demand (rho init)
  unsigned long long invaug_fx;   // fixed-point (2^-32) sum of the inverse augmented demand
  unsigned long long disp_sum;
  unsigned long long disp_max;
  unsigned long long check_ovf;
  unsigned long long row_vio;
  unsigned long long site_vio;    // strict legality counts of lg_check (cells)
  unsigned long long fence_vio;
  unsigned long long core_vio;
  unsigned long long rail_vio;
  unsigned long long pin_short;   // technology counts of lg_check (cells; edge_vio sums both sides)
  unsigned long long pin_access;
  unsigned long long edge_vio;
  unsigned long long phase_ticks[16];  // per-phase wall-clock ticks (LG_PHASE_TIMERS builds)
};

enum : int32_t {
  kErrNone = 0,
  kErrPartitionTooDense = 1,  // a partition needs more than kMaxNpad sort slots
  kErrUsageRange = 2,         // a pixel usage above 65535 (16-bit LDS packing)
};

constexpr int kWave = 64;
// waves per partition workgroup of the register-sort launches: the candidate
// moves of a cell are split over them (one per lane); the LDS-sort launches
// keep one wave
constexpr int kPartWaves = 4;
constexpr int kNumDirDefault = 245;
constexpr int kNumDirPadded = 256;
constexpr int kNumDirWide = 4093;
constexpr int kNumDirWidePadded = 4096;
// sort slots of a partition: pow2(cells) up to the reference's limit
// (assert(array_size_padded < 12288), CudaAdmmLegalizer.cu:709-710, lets a
// power of two reach 8192); up to kMaxNpadLds the sort buffers live in LDS,
// above it in a global-memory buffer of the partition's own
constexpr int kMaxNpad = 8192;
constexpr int kMaxNpadLds = 4096;
constexpr int kMaxStagnationWindow = 64;  // Ctl::ovf_window entries
constexpr int kBucketCells = 1024;  // cells per workgroup of k_bucket
constexpr int kNpadSmall = 1024;
constexpr int kLargeGroups = 128;  // workgroups per slot of the batched large-class launch

}  // namespace lgdev
