// ppr_query_exec: C = A * B over bit-tiled grids, B read from the transposed grid
// (the task core/task/gpu_task/ppr_query.cu runs; the reference's docs/README.md:54
// names this executable). Flags follow the reference apps (apps/gemm.cpp:17-21):
//   ppr_query_exec -i <A grid/> -it <A^T grid/> [-o <C grid/>] [-count N]
// Paths must end in '/'. Exit status 0 on success, 1 on any failure (message on stderr).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "host/matrixgraph_host.h"

namespace {
bool take(int& i, int argc, char** argv, const char* name, std::string* out) {
  const std::string a = argv[i];
  const std::string n1 = std::string("-") + name, n2 = std::string("--") + name;
  for (const std::string& n : {n1, n2}) {
    if (a == n && i + 1 < argc) {
      *out = argv[++i];
      return true;
    }
    if (a.rfind(n + "=", 0) == 0) {
      *out = a.substr(n.size() + 1);
      return true;
    }
  }
  return false;
}
}  // namespace

int main(int argc, char** argv) {
  std::string in, in_t, out, count = "0";
  for (int i = 1; i < argc; ++i) {
    if (take(i, argc, argv, "i", &in) || take(i, argc, argv, "it", &in_t) || take(i, argc, argv, "o", &out) ||
        take(i, argc, argv, "count", &count))
      continue;
    std::cerr << "unknown argument: " << argv[i] << "\nusage: " << argv[0]
              << " -i <grid/> -it <transposed grid/> [-o <output grid/>] [-count N]" << std::endl;
    return 1;
  }
  if (in.empty() || in_t.empty()) {
    std::cerr << "ppr_query_exec: -i and -it are required" << std::endl;
    return 1;
  }
  try {
    using namespace sics::matrixgraph::core;
    MatrixGraph sys(kCHBL);
    task::PPRQuery q(in, in_t, out, std::strtoull(count.c_str(), nullptr, 10));
    sys.Run(common::kPPRQuery, &q);
    std::cout << "tile_products: " << q.GetTileProducts() << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "ppr_query_exec: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
