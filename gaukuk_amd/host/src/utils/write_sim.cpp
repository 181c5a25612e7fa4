// write_sim.cpp -- output files in the reference's binary layout (src/utils/write_sim.cpp:14-58),
// read by its ReadGaukuk (pypkg/read_gaukuk.py) and by gaukuk_amd.gkio:
//   int32 sizeof(Real); Real t; int32 NVar, nx, ny, nz, nGhost; Real xc[lenx], yc[leny], zc[lenz];
//   int32 size; Real data[NVar][lenz][leny][lenx] (all ghost cells included).
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "../sim.hpp"

namespace Gaukuk {

void Sim::WriteData(const int outputID, DataType dType) {
    // a boundary update whose in-place part runs on the device is applied before writing
    if (boundary.DeviceUpdatePending()) EnsureDevice();
    if (hostStale_) {
        Check(gk_download_cons(ctx_, cons.data()), "gk_download_cons");
        hostStale_ = false;
    }
    std::ostringstream name;
    name << (dType == DataType::Prim ? "prim_" : "cons_") << std::setw(5) << std::setfill('0') << outputID;
    std::ofstream out(name.str(), std::ios::binary);
    if (!out) throw std::runtime_error("Failed to open " + name.str());
    const int realSize = sizeof(Real);
    const int hdr[5] = {NVar, grid.nx, grid.ny, grid.nz, grid.nGhost};
    out.write(reinterpret_cast<const char*>(&realSize), sizeof(int));
    out.write(reinterpret_cast<const char*>(&t), sizeof(Real));
    out.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
    out.write(reinterpret_cast<const char*>(domain.xc.data()), domain.xc.GetSizeInBytes());
    out.write(reinterpret_cast<const char*>(domain.yc.data()), domain.yc.GetSizeInBytes());
    out.write(reinterpret_cast<const char*>(domain.zc.data()), domain.zc.GetSizeInBytes());
    const TArray<Real>* src = &cons;
    if (dType == DataType::Prim) {
        eos.ConsToPrim(cons, prim, grid);
        src = &prim;
    }
    const int size = (int)src->GetSize();
    out.write(reinterpret_cast<const char*>(&size), sizeof(int));
    out.write(reinterpret_cast<const char*>(src->data()), src->GetSizeInBytes());
    if (!out) throw std::runtime_error("Failed to write " + name.str());
}

}  // namespace Gaukuk
