/**
 * @file utils.hh
 * Host helpers of the reference utils.hh:14-73: compensated (Kahan) sums of a vector or of
 * one EventVector column, a uniqueness check and mean/rms.
 */
#ifndef MOREFIT_UTILS_HH
#define MOREFIT_UTILS_HH

#include <cmath>
#include <set>
#include <string>
#include <vector>

#include "eventvector.hh"

namespace morefit {

template <typename returnT = double, typename vectorT = double>
returnT kahan_summation(const std::vector<vectorT>& values) {
  returnT sum = 0.0, c = 0.0;
  for (vectorT v : values) {
    returnT y = (returnT)v - c;
    volatile returnT t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

template <typename returnT = double, typename vectorT = double>
returnT kahan_summation(const EventVector<vectorT, returnT>& data, int dim = 0) {
  returnT sum = 0.0, c = 0.0;
  for (std::size_t i = 0; i < data.nevents(); i++) {
    returnT y = (returnT)data(i, (unsigned)dim) - c;
    volatile returnT t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

inline bool elements_unique(const std::vector<std::string>& v) { return std::set<std::string>(v.begin(), v.end()).size() == v.size(); }

template <typename T = double, typename U = double>
void mean_rms(const std::vector<T>& v, U& mean, U& rms) {
  mean = 0.0;
  rms = 0.0;
  if (v.empty()) return;
  for (auto x : v) mean += x;
  mean /= v.size();
  for (auto x : v) rms += (x - mean) * (x - mean);
  rms = v.size() > 1 ? std::sqrt(rms / (v.size() - 1)) : 0.0;
}

}  // namespace morefit
#endif
