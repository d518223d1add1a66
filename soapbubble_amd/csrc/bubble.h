// C++ host surface of the soap-film step over the C ABI of libsoapfilm.so:
// the reference's Solver (include/solver.cuh:107-150) and Bubble
// (include/bubble.cuh:36-89, kernel/bubble.cu:2566-2704) classes with their
// constructor arguments, stepping calls, texture initialisation and
// thickness output, and the config.txt reader of kernel/main.cu:4-56.
//
// Errors throw sfb::Error (the reference exits the process instead).
#pragma once

#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/soapfilm.h"

namespace sfb {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// config.txt, read positionally as kernel/main.cu:14-37 does: one
// "name value" pair per entry in this order.
struct Config {
    double radius = 0, H = 0, U = 1, c_m = 0, Gamma_m = 0, T = 0, Ds = 0, rm = 0;
    int nTheta = 0;
    double dt = 0, DT = 0;    // timePerStep, timePerFrame (s)
    int frames = 0;
    std::string outputDir, thicknessImage, AMGconfig;
    double blendCoeff = 0.5;
    int device = 0;
};
// Returns false and a message on a missing file or a malformed entry.
bool read_config(const std::string& path, Config& out, std::string& err);

// The outer-solver settings of an AmgX JSON configuration (init/PCGF_F.json):
// "tolerance", "convergence" (ABSOLUTE or RELATIVE...) and "max_iters" of the
// top-level "solver" object.
struct SolverSettings {
    double tolerance = 1e-7;
    bool absolute = true;
    int max_iters = 100;
};
bool read_amg_settings(const std::string& path, SolverSettings& out, std::string& err);

class Solver {
public:
    // Solver::Solver (solver.cu:63-160): nPhi = 2 nTheta; dt in seconds.
    // Gamma_m and T (config.txt) reach the reference's kernels as constants
    // uploaded by Bubble::run; here they are constructor arguments.
    // The scalar advection is BIMOCQ with five samples per point, as
    // header.cuh:97-98 builds the reference, with blendCoeff `blend`;
    // bimocq = false selects the plain advectionAllCentered.
    // viscous = true selects the VISCOUS stress path (header.cuh:99, off in
    // the reference).
    Solver(size_t nPhi, size_t nTheta, double radius, double dt, double H, int device,
           const std::string& AMGconfig, sf_precision precision = SF_PREC_FP64,
           double Gamma_m = 1e-6, double T = 298.15, bool bimocq = true, double blend = 0.5,
           bool viscous = false);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    void initThicknessfromPic(const std::string& path);   // initializer.cu:50-94
    void initAirflowfromPic(const std::string& path);     // initializer.cu:97-135
    void initWithConst(sf_field field, double val);       // initializer.cu:22-31
    void stepForward();                                   // solver.cu:260-292
    void stepForward(double dt);                          // solver.cu:252-257, dt in s
    // n steps of stepForward() as one call: a solve that misses its
    // tolerance stops the bubble on the device (its later steps are no-ops),
    // so this equals the reference's loop of stepForward() guarded by
    // isBroken() (bubble.cu:2669), and the steps replay as captured graphs
    void stepForward(int n);
    // the broken flag of write_thickness_img (no cell with eta > 0 at the
    // last frame written, solver.cu:568-575) or a solve that missed its
    // tolerance (AlgebraicMultiGridCG, bubble.cu:1500-1507)
    bool isBroken();                                      // solver.cu:296-298
    void setBroken(bool broken) { broken_ = broken; }
    void setTime(double t);                               // currentTimeGlobal (s)
    // write_thickness_img (solver.cu:541-593): the frame's pixels are
    // converted on the device, and the file is written on a thread of its
    // own while the caller goes on stepping (at most kWriters files in
    // flight: the oldest write is finished, and its failure reported,
    // before its buffer takes a new frame); finish_writes() and the
    // destructor finish them all
    void write_thickness_img(const std::string& s, int frame);
    void finish_writes();
    void copyToCPU(sf_field field, std::vector<double>& out);    // solver.cu:243-249
    sf_ctx* context() { return ctx_; }
    size_t nPhi() const { return nPhi_; }
    size_t nTheta() const { return nTheta_; }

private:
    void check(int rc);
    sf_ctx* ctx_ = nullptr;
    size_t nPhi_, nTheta_;
    double radius_, dt_;
    bool broken_ = false;
    // the last statistics read back, valid until the next step or upload
    // (the broken checks of a frame share one read)
    sf_bubble_stats stats_{};
    bool stats_fresh_ = false;
    const sf_bubble_stats& stats();
    static constexpr int kWriters = 2;
    struct Writer {
        std::vector<uint16_t> frame;   // the frame being written
        std::thread thread;
        std::string error;
    };
    Writer writers_[kWriters];
    int next_writer_ = 0;
    void finish(Writer& w);
};

class Bubble {
public:
    // Bubble::Bubble (bubble.cu:2568-2580); dt and DT in seconds.
    Bubble(double radius, double H, double U, double c_m, double Gamma_m, double T, double Ds,
           double rm, size_t nTheta, double dt, double DT, int frames, std::string outputDir,
           std::string thicknessImage, int device, std::string AMGconfig, double blendCoeff);
    explicit Bubble(const Config& c);
    // The air texture initAirflowfromPic reads (bubble.cu:2641 uses
    // ../init/velocityfield_512.exr relative to the build directory).
    std::string airImage = "../init/velocityfield_512.exr";
    sf_precision precision = SF_PREC_FP64;
    bool bimocq = true;   // header.cuh:97 defines BIMOCQ
    bool viscous = false; // header.cuh:99 leaves VISCOUS undefined
    bool verbose = true;
    // Bubble::run (bubble.cu:2609-2704): frames of DT seconds, each made of
    // steps of dt and a partial last step; the air-flow time is the frame
    // start; stops when the film breaks. Writes frame%04d.exr into
    // outputDir (frame 0 is the initial state). Returns the frames completed.
    int run();
    // wall time of the last run()'s frame loop (frames 1.. and the writes
    // they started), s: the frames' cost without the process start,
    // textures and context set-up
    double loopSeconds() const { return loop_seconds_; }

private:
    double loop_seconds_ = 0.0;
    double radius_, H_, Gamma_m_, T_, dt_, DT_, blend_;
    size_t nTheta_;
    int frames_, device_;
    std::string outputDir_, thicknessImage_, AMGconfig_;
};

}  // namespace sfb
