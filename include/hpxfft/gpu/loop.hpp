// loop.hpp -- hpxfft::gpu::loop: the MI355X drop-in for HPX-FFT's
// shared-memory 2-D r2c class hpxfft::shared::loop
// (core/include/hpxfft/shared/loop.hpp:15-57, core/src/shared/loop.cpp).
//
// Header-only C++ layer over the C ABI of include/hpxfft_hip.h; link with
// -lhpxfft_hip.  Member signatures are identical to the reference's:
//   void initialize(vector_2d values_vec, const std::string PLAN_FLAG);  // loop.hpp:20
//   vector_2d fft_2d_r2c_par();  vector_2d fft_2d_r2c_seq();              // loop.hpp:22,24
//   real get_measurement(std::string name);                               // loop.hpp:26
//   void write_plans_to_file(std::string file_path);                      // loop.hpp:28
// plus vector_2d fft_2d_r2c() so that the task-graph variants sync/opt/naive
// (sync.hpp:23-27, opt.hpp:23-27, naive.hpp:23-27) can be the same class.
//
// Protocol (as in the reference): initialize() takes the buffer by value
// (callers std::move it in), fft_2d_r2c*() returns it transformed and leaves
// the object without data; call initialize() again before the next transform
// (a second call throws std::runtime_error instead of reading a moved-from
// buffer, loop.cpp:112).  Errors: a bad plan flag or an unsupported size
// throws std::invalid_argument (adapter_fftw.hpp:42); device, allocation and
// file errors throw std::runtime_error (loop.cpp:200).  Unknown measurement
// keys read 0 and are inserted (std::map::operator[], loop.cpp:192).
#ifndef HPXFFT_GPU_LOOP_HPP
#define HPXFFT_GPU_LOOP_HPP

#include "../../hpxfft_hip.h"
#include "../util/vector_2d.hpp"

#include <map>
#include <stdexcept>
#include <string>
#include <utility>

typedef double real;   // as the reference's core/include/hpxfft/shared/loop.hpp:9

namespace hpxfft::gpu
{
using vector_2d = hpxfft::util::vector_2d<real>;

inline void throw_on_status(int status)
{
    if (status == HPXFFT_OK) return;
    const std::string msg = hpxfft_last_error();
    if (status == HPXFFT_INVALID_ARG || status == HPXFFT_UNSUPPORTED_SIZE) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

struct loop
{
  public:
    loop() = default;
    loop(const loop &) = delete;
    loop &operator=(const loop &) = delete;
    ~loop() { release(); }

    void initialize(vector_2d values_vec, const std::string PLAN_FLAG)
    {
        values_vec_ = std::move(values_vec);
        dim_c_x_ = values_vec_.n_row();
        dim_c_y_ = values_vec_.n_col() / 2;
        dim_r_y_ = 2 * dim_c_y_ - 2;
        int effort = 0;
        throw_on_status(hpxfft_plan_flag_from_string(PLAN_FLAG.c_str(), &effort));
        if (values_vec_.n_col() % 2 != 0 || dim_c_y_ < 2)
            throw std::invalid_argument("n_col must be even and >= 4 (Ny = n_col - 2)");
        release();
        throw_on_status(hpxfft_plan_r2c_2d(&plan_, dim_c_x_, dim_r_y_, values_vec_.n_col(), HPXFFT_F64, effort, nullptr));
        double v = 0.0;
        throw_on_status(hpxfft_get_measurement(plan_, "plan", &v));
        measurements_["plan"] = v;
        throw_on_status(hpxfft_get_measurement(plan_, "plan_flops", &v));
        measurements_["plan_flops"] = v;
        ready_ = true;
    }

    vector_2d fft_2d_r2c_par() { return run(); }
    vector_2d fft_2d_r2c_seq() { return run(); }
    vector_2d fft_2d_r2c() { return run(); }

    real get_measurement(std::string name) { return measurements_[name]; }

    void write_plans_to_file(std::string file_path)
    {
        if (!plan_) throw std::runtime_error("write_plans_to_file called before initialize");
        throw_on_status(hpxfft_write_plan(plan_, file_path.c_str()));
    }

  private:
    vector_2d run()
    {
        if (!ready_ || !plan_) throw std::runtime_error("fft_2d_r2c called without a preceding initialize");
        throw_on_status(hpxfft_exec_r2c_2d_host(plan_, values_vec_.data()));
        static const char *keys[] = {"total", "first_fftw", "first_trans", "second_fftw", "second_trans", "h2d", "d2h"};
        for (const char *k : keys)
        {
            double v = 0.0;
            throw_on_status(hpxfft_get_measurement(plan_, k, &v));
            measurements_[k] = v;
        }
        ready_ = false;
        return std::move(values_vec_);
    }

    void release()
    {
        if (plan_) hpxfft_plan_destroy(plan_);
        plan_ = nullptr;
    }

    std::size_t dim_r_y_ = 0, dim_c_y_ = 0, dim_c_x_ = 0;
    hpxfft_plan plan_ = nullptr;
    vector_2d values_vec_;
    std::map<std::string, real> measurements_;
    bool ready_ = false;
};

// the shared-memory task-graph variants compute the same transform
using sync = loop;
using opt = loop;
using naive = loop;
}  // namespace hpxfft::gpu

#endif  // HPXFFT_GPU_LOOP_HPP
