/* Shared part of the oracle (TEST INFRASTRUCTURE ONLY, see oracle.h). */
#include <math.h>

#include "oracle.h"

#define ORC_PI 3.14159265358979323846

/* ElementsData::updateWavelet / Ricker / Berlage (fem_utils/src/fem_utils.cpp:53-82).
 * All intermediate values are float as in the reference; M_PI products are
 * formed in double and rounded to float on assignment, as there. */
float orc_wavelet_value(double td, const orc_wavelet *w) {
  float t = (float)td;
  float ts = (float)w->timeshift, freq = (float)w->freq, ampl = (float)w->ampl;
  if (t < ts) return 0.f;
  t = t - ts;
  if (w->type == 1) {
    float tmp = (float)(ORC_PI * (double)freq * (double)t - ORC_PI);
    return ampl * (1.f - 2.f * tmp * tmp) * expf(-1.f * tmp * tmp);
  }
  if (w->type == 2) {
    float w0 = (float)(2 * ORC_PI * (double)freq);
    float w1 = w0 / sqrtf(3.f);
    return ampl * w1 * w1 / 4.f * expf(-1.f * w1 * t) *
           (sinf(w0 * t) * (1.f / (w1 * w1 * w1) + t / (w1 * w1) - t * t / w1) -
            cosf(w0 * t) * sqrtf(3.f) * (t * t / w1 + t / (w1 * w1)));
  }
  return 0.f;
}

/* The same wavelets evaluated in double (the FP64 restatement). */
double orc_wavelet_value_f64(double t, const orc_wavelet *w) {
  double ts = w->timeshift, freq = w->freq, ampl = w->ampl;
  if (t < ts) return 0.0;
  t = t - ts;
  if (w->type == 1) {
    double tmp = ORC_PI * freq * t - ORC_PI;
    return ampl * (1.0 - 2.0 * tmp * tmp) * exp(-1.0 * tmp * tmp);
  }
  if (w->type == 2) {
    double w0 = 2 * ORC_PI * freq;
    double w1 = w0 / sqrt(3.0);
    return ampl * w1 * w1 / 4.0 * exp(-1.0 * w1 * t) *
           (sin(w0 * t) * (1.0 / (w1 * w1 * w1) + t / (w1 * w1) - t * t / w1) -
            cos(w0 * t) * sqrt(3.0) * (t * t / w1 + t / (w1 * w1)));
  }
  return 0.0;
}
