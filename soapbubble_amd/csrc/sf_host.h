// Host-only helpers of libsoapfilm.so (no device code).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/soapfilm.h"

namespace sfh {

// Bubble::Bubble / Bubble::run parameter derivation (bubble.cu:2573-2580,
// 2627-2638).
void derive_params(const sf_bubble_desc& d, sf_derived& o);

// Time-independent part of the air-flow corridor of airFlowU / airFlowV
// (bubble.cu:2317-2335, 2358-2375) projected on e_phi at u_phi faces (su,
// nth*nph) and on e_theta at u_theta faces (sv, rows 0..nth-2); the kernels
// multiply it by 3 - |t - 3|. NaN marks points where the reference's
// normalisation divides by zero.
void air_shapes(int nth, double radius, double U, double* su, double* sv);

// Real spherical harmonic, orthonormal on the unit sphere.
double real_ylm(int l, int m, double theta, double phi);

// Synthetic initial fields (see sf_synth_fields in soapfilm.h).
void synth_fields(int nth, uint64_t seed, int lmax, int lair, double noise, double air_amp,
                  double* eta, double* gam, double* uair0, double* vair0);

// HALF-float RGB OpenEXR, scanline, uncompressed, channels B,G,R
// (write_image, solver.cu:338-385); value eta * H * 2e5 in every channel.
bool write_thickness_exr(const std::string& path, int nth, int nph, const double* eta, double H);
// write_image_exr of a gray image given as its binary16 values
bool write_half_exr(const std::string& path, int w, int h, const uint16_t* gray);
// write_image (solver.cu:338-385): planes r, g, b of w x h floats as HALF
// channels B, G, R.
bool write_image_exr(const std::string& path, int w, int h, const float* r, const float* g,
                     const float* b);
// Reads one channel of an uncompressed HALF/FLOAT scanline EXR.
bool read_exr_channel(const std::string& path, const std::string& channel, int* width,
                      int* height, std::vector<float>* data);

// initThicknessfromPic / initAirflowfromPic (initializer.cu:50-135): the
// B, G, R channels of an EXR texture resized to nPhi x nTheta by bilinear
// interpolation with pixel centres aligned and replicated borders (the
// semantics of OpenCV's INTER_LINEAR resize, which the reference calls);
// eta = (B - 0.5) 0.3 + 0.5; u_air = (G_i + G_{i-1}) / 2 * radius on u_phi
// faces; v_air = -(R_j + R_{j+1}) / 2 * radius on u_theta faces. A null or
// empty path, or a file that cannot be read, leaves the output untouched
// and clears its bit of *loaded (bit 0 thickness, bit 1 air).
void texture_fields(const char* thickness_path, const char* air_path, int nth, double radius,
                    double* eta, double* uair0, double* vair0, int* loaded);

}  // namespace sfh
