// ORACLE -- test infrastructure only (see oracle.hpp).
// CPU identification backend: the reference's serial identification loops.
#pragma once

#include <cstdint>
#include <vector>

#include "../cctag_amd/csrc/host/ident.hpp"
#include "../cctag_amd/csrc/host/pool.hpp"

namespace oracle {

class CpuIdentBackend : public cctagx::IdentBackend {
 public:
  CpuIdentBackend(const uint8_t* img, int w, int h, std::vector<std::vector<float>> bank)
      : img_(img), w_(w), h_(h), bank_(std::move(bank)) {}
  void select_cuts(const std::vector<cctagx::CutJob>& jobs, const cctagx::HostParams& p,
                   std::vector<cctagx::CutSelection>& out) override;
  void refine(const std::vector<cctagx::RefineJob>& jobs, const cctagx::HostParams& p,
              std::vector<cctagx::RefineResult>& out) override;
  // Candidates run in parallel on `pool` (null: serially), as the reference
  // identifies them under tbb::parallel_for.
  void set_pool(cctagx::TaskPool* pool) { pool_ = pool; }

 private:
  cctagx::TaskPool* pool_ = nullptr;
  const uint8_t* img_;
  int w_, h_;
  std::vector<std::vector<float>> bank_;
};

}  // namespace oracle
