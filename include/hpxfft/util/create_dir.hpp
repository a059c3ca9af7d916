// create_dir.hpp -- hpxfft::util::create_parent_dir with the API of HPX-FFT's
// core/include/hpxfft/util/create_dir.hpp:9 (defined in
// core/src/util/create_dir.cpp:3-18), header-only here so that the
// reference's harness links against -lhpxfft_hip alone.
//
// Creates the directory that will hold `file_path` (with its parents) when
// it does not exist, announcing it on stdout as "Directory created: <dir>",
// and throws std::runtime_error("Failed to create directory: <dir>") when it
// cannot be made.  A path with no parent part names a file in the working
// directory, which exists: nothing is created (the reference would try to
// create the empty path and throw).
#ifndef HPXFFT_GPU_CREATE_DIR_HPP
#define HPXFFT_GPU_CREATE_DIR_HPP

#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <system_error>

namespace hpxfft::util
{
inline void create_parent_dir(std::filesystem::path file_path)
{
    const std::filesystem::path dir = file_path.parent_path();
    if (dir.empty() || std::filesystem::exists(dir)) return;
    std::error_code ec;
    if (!std::filesystem::create_directories(dir, ec) || ec)
        throw std::runtime_error("Failed to create directory: " + dir.string());
    std::cout << "Directory created: " << dir << "\n";
}
}  // namespace hpxfft::util

#endif
