// report_format.cpp -- prints report lines written by tapas_gpu::util::RankCSV
// (no GPU): the Python test compares them with the reference's RankCSV format
// (cpp/include/tapas/util.h:114-317).
#include <cstdio>

#include "../../tapas-reduce_amd/tapas_gpu/tapas_fmm.h"

int main() {
  tapas_gpu::util::RankCSV a({"total", "numP2P", "Hostname"});
  a.At("total", 1.5);
  a.At("numP2P", (uint64_t)123);
  a.At("Hostname", std::string("a-very-long-hostname"));
  fputs(a.Text().c_str(), stdout);
  tapas_gpu::util::RankCSV b({"LET-Res-acomm", "grow-global"});
  b.At("LET-Res-acomm", 0.0);
  b.At("grow-global", -2.5e-7);
  fputs(b.Text().c_str(), stdout);
  return 0;
}
