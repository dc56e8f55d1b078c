/**
 * @file dimensionvector.hh
 * Observables ("dimensions") with a range.  API of the reference dimensionvector.hh:22-68
 * (name, tex label, min/max, unlimited flag, get_from_name/get_to_name).
 */
#ifndef MOREFIT_DIMENSIONVECTOR_HH
#define MOREFIT_DIMENSIONVECTOR_HH

#include <algorithm>
#include <limits>
#include <string>
#include <vector>

namespace morefit {

template <typename T = double>
class dimension {
 public:
  std::string name_;
  std::string tex_;
  T min_;
  T max_;
  bool unlimited_;
  explicit dimension(std::string name)
      : name_(name), tex_(name), min_(-std::numeric_limits<T>::infinity()), max_(std::numeric_limits<T>::infinity()), unlimited_(true) {}
  dimension(std::string name, T min, T max, bool unlimited = false)
      : name_(name), tex_(name), min_(min), max_(max), unlimited_(unlimited) {}
  dimension(std::string name, std::string tex, T min, T max, bool unlimited = false)
      : name_(name), tex_(tex), min_(min), max_(max), unlimited_(unlimited) {}
  std::string get_name() const { return name_; }
  std::string get_tex() const { return tex_; }
  // names of the integration-boundary variables used by definite integrals
  std::string get_from_name() const { return "morefit_" + name_ + "_from"; }
  std::string get_to_name() const { return "morefit_" + name_ + "_to"; }
  T get_min() const { return min_; }
  T get_max() const { return max_; }
  // centre of the range (0 for unlimited observables); FP32 kernels store offsets from it
  T get_centre() const { return unlimited_ ? T(0) : T(0.5) * (min_ + max_); }
  bool is_unlimited() const { return unlimited_; }
  void set_name(std::string name) { name_ = name; }
  void set_tex(std::string tex) { tex_ = tex; }
  void set_min(T min) { min_ = min; }
  void set_max(T max) { max_ = max; }
  void set_unlimited(bool unlimited) { unlimited_ = unlimited; }
};

template <typename T = double>
class DimensionVector {
  std::vector<dimension<T>> dimensions_;

 public:
  explicit DimensionVector(std::vector<dimension<T>> dimensions) : dimensions_(dimensions) {}
  dimension<T> get_dimension(int idx) const { return dimensions_.at(idx); }
  dimension<T> get_dimension(std::string name) const {
    auto it = std::find_if(dimensions_.begin(), dimensions_.end(), [&name](const dimension<T>& d) { return d.get_name() == name; });
    return *it;
  }
  int size() const { return (int)dimensions_.size(); }
  void clear() { dimensions_.clear(); }
  void push_back(const dimension<T>& d) { dimensions_.push_back(d); }
  void push_back(const DimensionVector<T>& v) { dimensions_.insert(dimensions_.end(), v.dimensions_.begin(), v.dimensions_.end()); }
};

}  // namespace morefit
#endif
