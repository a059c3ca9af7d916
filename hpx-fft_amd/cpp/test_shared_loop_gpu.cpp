// test_shared_loop_gpu -- the reference's known-answer test for the shared
// classes (test/src/test_shared_loop.cpp:15-53 and its siblings
// test_shared_sync.cpp, test_shared_opt.cpp, test_naive.cpp,
// test_shared_agas.cpp) run against hpxfft::shared::{loop,sync,opt,naive}
// resolved to the MI355X class, with the same exact (bitwise) comparison.
// Also checks the error behaviour of the surface.  Plain asserts (no
// Catch2); exit status 0 = all passed.  Needs a HIP device.
#include "hpxfft/shared/loop.hpp"
#include "hpxfft/shared/naive.hpp"
#include "hpxfft/shared/opt.hpp"
#include "hpxfft/shared/sync.hpp"

#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>

static int failures = 0;
#define CHECK(cond)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            std::cerr << "FAILED: " #cond " (line " << __LINE__ << ")\n";                                             \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

static hpxfft::shared::vector_2d input_4x6()
{
    hpxfft::shared::vector_2d values_vec(4, 6, 0.0);
    for (std::size_t i = 0; i < 4; ++i)
    {
        values_vec(i, 0) = 1.0;
        values_vec(i, 1) = 2.0;
        values_vec(i, 2) = 3.0;
        values_vec(i, 3) = 4.0;
    }
    return values_vec;
}

static hpxfft::shared::vector_2d expected_4x6()
{
    hpxfft::shared::vector_2d e(4, 6, 0.0);
    e(0, 0) = 40.0;
    e(0, 2) = -8.0;
    e(0, 3) = 8.0;
    e(0, 4) = -8.0;
    return e;
}

template <typename C>
static void known_answer(const char *name, const std::string &flag)
{
    C fft;
    fft.initialize(input_4x6(), flag);
    hpxfft::shared::vector_2d out = fft.fft_2d_r2c();
    const double total = fft.get_measurement("total");
    CHECK(total >= 0.0);
    CHECK(fft.get_measurement("plan_flops") > 0.0);
    CHECK(fft.get_measurement("no such key") == 0.0);
    const bool ok = out == expected_4x6();
    CHECK(ok);
    std::cout << name << " (" << flag << "): " << (ok ? "exact match" : "MISMATCH") << std::endl;
}

static int run_all();

int main()
{
    try
    {
        return run_all();
    }
    catch (const std::exception &e)
    {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 3;
    }
}

static int run_all()
{
    // loop, as test_shared_loop.cpp:46-53 (fft_2d_r2c_par, "estimate")
    {
        hpxfft::shared::loop fft2;
        std::string plan_flag = "estimate";
        fft2.initialize(input_4x6(), plan_flag);
        hpxfft::shared::vector_2d out2 = fft2.fft_2d_r2c_par();
        CHECK(fft2.get_measurement(std::string("total")) >= 0.0);
        CHECK(out2 == expected_4x6());
        // the buffer was moved out: a second transform without initialize is an error
        bool threw = false;
        try
        {
            (void)fft2.fft_2d_r2c_par();
        }
        catch (const std::runtime_error &)
        {
            threw = true;
        }
        CHECK(threw);
        // re-initialize and use the sequential entry point
        fft2.initialize(input_4x6(), plan_flag);
        CHECK(fft2.fft_2d_r2c_seq() == expected_4x6());
    }
    known_answer<hpxfft::shared::sync>("sync", "estimate");
    known_answer<hpxfft::shared::opt>("opt", "estimate");
    known_answer<hpxfft::shared::naive>("naive", "estimate");
    known_answer<hpxfft::shared::loop>("loop", "measure");   // test_shared_agas.cpp:38 plans with "measure"

    // invalid plan flag -> std::invalid_argument (adapter_fftw.hpp:42)
    {
        bool threw = false;
        hpxfft::shared::loop f;
        try
        {
            f.initialize(input_4x6(), "fastest");
        }
        catch (const std::invalid_argument &e)
        {
            threw = std::string(e.what()) == "Invalid FFTW plan flag string";
        }
        CHECK(threw);
    }
    // unsupported size -> std::invalid_argument
    {
        bool threw = false;
        hpxfft::shared::loop f;
        try
        {
            // Ny = 8388638: Ny/2 = 4194319 is a prime above 2^22, the longest length
            // Bluestein's algorithm takes over two levels
            f.initialize(hpxfft::shared::vector_2d(1, 8388640, 1.0), "estimate");
        }
        catch (const std::invalid_argument &)
        {
            threw = true;
        }
        CHECK(threw);
    }
    // unwritable plan file -> std::runtime_error (loop.cpp:198-201)
    {
        bool threw = false;
        hpxfft::shared::loop f;
        f.initialize(input_4x6(), "estimate");
        try
        {
            f.write_plans_to_file("/nonexistent-dir/plan.txt");
        }
        catch (const std::runtime_error &e)
        {
            threw = std::string(e.what()).rfind("Failed to open file: ", 0) == 0;
        }
        CHECK(threw);
    }
    std::cout << (failures ? "FAILED" : "ALL PASSED") << std::endl;
    return failures ? 1 : 0;
}
