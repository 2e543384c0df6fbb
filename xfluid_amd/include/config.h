// config.h -- frame output location, as the reference's config.h:4-7
// (Config::outputFild); meshes are written as <outputFild><frame>.ply.
#pragma once
#include <string>

namespace Config {
inline std::string outputFild = "./xfluid_output/";
}
