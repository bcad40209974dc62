// Exercises the C++ host API (include/hemepure_amd.hpp) the way reference
// code would: LatticeData accessors, then (mode "run") an LBM with TRT + BFL
// selected by type and a SimulationMaster run, or (mode "extract") a run
// writing a checkpoint file and a plane of pressure and velocity, then a
// second run restarted from the checkpoint.  Prints one line per fact for the
// test to compare with the Python binding.
#include <cstdio>
#include <cstring>
#include <string>

#include "hemepure_amd.hpp"

using namespace hemelb;

typedef lb::lattices::D3Q19 Lat;
typedef lb::kernels::TRT<Lat> Kernel;
typedef lb::LBM<Lat, Kernel, lb::streamers::BouzidiFirdaousLallemand<Kernel>,
                lb::streamers::NashZerothOrderPressureIolet<Kernel>>
    TrtBflNash;

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "tables";
  try {
    geometry::Geometry g = geometry::Geometry::Cylinder(6.0, 24, 16);
    geometry::LatticeData ld(g);
    std::printf("sites %lld\n", (long long)ld.GetLocalFluidSiteCount());
    for (unsigned c = 0; c < 6; ++c) std::printf("class %u %lld\n", c, (long long)ld.GetMidDomainCollisionCount(c));
    const site_t k = ld.GetLocalFluidSiteCount() - 1;
    const auto x = ld.GetGlobalSiteCoords(k);
    std::printf("last %lld %lld %lld type %d ctype %u\n", (long long)x.x, (long long)x.y, (long long)x.z,
                ld.GetSiteType(k), ld.GetCollisionType(k));
    long long pushSum = 0;
    for (site_t i = 0; i < ld.GetLocalFluidSiteCount(); ++i)
      for (unsigned d = 0; d < 19; ++d) pushSum += ld.GetStreamedIndex(i, d);
    std::printf("pushsum %lld\n", pushSum);
    if (mode == "tables") return 0;

    const auto io = g.Iolets();
    double pin[3], nin[3], pout[3], nout[3];
    for (int a = 0; a < 3; ++a) {
      pin[a] = io[0].position[a];
      nin[a] = io[0].normal[a];
      pout[a] = io[1].position[a];
      nout[a] = io[1].normal[a];
    }
    SimulationMaster<TrtBflNash>::Config cfg{g, lb::LbmParameters(0.7), {lb::iolets::InOutLetCosine(pin, nin, 1.002)},
                                             {lb::iolets::InOutLetCosine(pout, nout, 1.0)}};
    if (mode == "extract") {
      const std::string dir = argv[2];
      cfg.totalSteps = 150;
      cfg.voxelSize = 1e-4;
      cfg.timeStepLength = 1e-5;
      extraction::PropertyOutputFile cp;
      cp.filename = dir + "/cp.xtr";
      cp.frequency = 50;
      cp.fields = {{extraction::OutputField::Distributions, ""}};
      extraction::PropertyOutputFile flow;
      flow.filename = dir + "/flow.xtr";
      flow.frequency = 75;
      flow.fields = {{extraction::OutputField::Pressure, "p"}, {extraction::OutputField::Velocity, ""}};
      flow.selector = HLB_SELECT_PLANE;
      flow.point[2] = 12 * cfg.voxelSize;
      cfg.propertyOutputs = {cp, flow};
      {
        SimulationMaster<TrtBflNash> master(cfg);
        master.RunSimulation();
        std::printf("timestep %lld\n", (long long)master.GetLBM().GetTimeStep());
      }
      auto again = cfg;
      again.propertyOutputs.clear();
      again.checkpointFile = dir + "/cp.xtr";
      again.checkpointTime = 100;
      SimulationMaster<TrtBflNash> restarted(again);
      std::printf("restart %lld\n", (long long)restarted.GetLBM().GetTimeStep());
      restarted.RunSimulation();
      std::printf("final %lld\n", (long long)restarted.GetLBM().GetTimeStep());
      const auto f = restarted.GetLBM().GetDistributions();
      FILE* out = std::fopen((dir + "/restart.bin").c_str(), "wb");
      if (!out) return 4;
      std::fwrite(f.data(), sizeof(double), f.size(), out);
      std::fclose(out);
      return 0;
    }
    cfg.totalSteps = 150;
    cfg.propertyPeriod = 150;
    cfg.stabilityPeriod = 50;
    SimulationMaster<TrtBflNash> master(cfg);
    master.RunSimulation();
    auto& lbm = master.GetLBM();
    const auto& cache = lbm.GetPropertyCache();
    std::printf("timestep %lld\n", (long long)lbm.GetTimeStep());
    const auto f = lbm.GetDistributions();
    if (argc > 2) {  // rho, u, f as raw doubles for a bitwise comparison
      FILE* out = std::fopen(argv[2], "wb");
      if (!out) return 4;
      std::fwrite(cache.densityCache.data(), sizeof(double), cache.densityCache.size(), out);
      std::fwrite(cache.velocityCache.data(), sizeof(double), cache.velocityCache.size(), out);
      std::fwrite(f.data(), sizeof(double), f.size(), out);
      std::fclose(out);
    }
    std::printf("stable %d\n", lbm.IsStable() ? 1 : 0);
  } catch (const Exception& e) {
    std::printf("error %s\n", e.what());
    return 3;
  }
  return 0;
}
