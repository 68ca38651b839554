// spatial_grid_tests.cpp -- the reference's broad-phase test
// (ball2dtests/collision_detection_tests.cpp:18-74) written against the C++ host
// interface of the MI355X library: the grid result must equal the all-pairs result.
// The AABB fixtures (the reference's test_00/01/02 data) are read from a raw file of
// doubles [minx, miny, maxx, maxy] per box, written by the pytest driver.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "scx_scisim.hpp"

using scx_scisim::AABB;

static int executeSpatialGridTest(const std::vector<double>& aabb_data) {
  if (aabb_data.size() % 4 != 0) return EXIT_FAILURE;
  const unsigned num_aabbs = static_cast<unsigned>(aabb_data.size() / 4);
  std::vector<AABB<2>> aabbs(num_aabbs);
  for (unsigned k = 0; k < num_aabbs; ++k) {
    aabbs[k].min() = {aabb_data[4 * k + 0], aabb_data[4 * k + 1]};
    aabbs[k].max() = {aabb_data[4 * k + 2], aabb_data[4 * k + 3]};
  }
  scx_scisim::Device dev(2, num_aabbs);
  std::set<std::pair<unsigned, unsigned>> spatial_grid_overlaps;
  scx_scisim::SpatialGridDetector::getPotentialOverlaps(dev, aabbs, spatial_grid_overlaps);
  std::set<std::pair<unsigned, unsigned>> all_pairs_overlaps;
  scx_scisim::SpatialGridDetector::getPotentialOverlapsAllPairs(aabbs, all_pairs_overlaps);
  const bool active_sets_agree = spatial_grid_overlaps == all_pairs_overlaps;
  std::cout << "Num AABBs:               " << num_aabbs << std::endl;
  std::cout << "Collisions spatial grid: " << spatial_grid_overlaps.size() << std::endl;
  std::cout << "Collisions all pairs:    " << all_pairs_overlaps.size() << std::endl;
  std::cout << "Active sets agree:       " << active_sets_agree << std::endl;
  return active_sets_agree ? EXIT_SUCCESS : EXIT_FAILURE;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "Usage: " << argv[0] << " fixture.bin" << std::endl;
    return EXIT_FAILURE;
  }
  std::ifstream in(argv[1], std::ios::binary);
  std::vector<double> data;
  double x;
  while (in.read(reinterpret_cast<char*>(&x), sizeof(double))) data.push_back(x);
  try {
    return executeSpatialGridTest(data);
  } catch (const scx_scisim::Error& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
