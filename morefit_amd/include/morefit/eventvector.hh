/**
 * @file eventvector.hh
 * Structure-of-arrays event container.  API and layout of the reference eventvector.hh:31-294:
 * element (event n, dimension d) lives at data[d * nevents_padded + n]; padding rounds the
 * per-column length up to a multiple of `padding` ((pad - n % pad) % pad extra slots,
 * eventvector.hh:109-119); an optional weight column (77-107).  Storage is a std::vector
 * and sizes are 64-bit, so columns above 4 GiB (large HBM-resident studies) are fine.
 */
#ifndef MOREFIT_EVENTVECTOR_HH
#define MOREFIT_EVENTVECTOR_HH

#include <algorithm>
#include <cstddef>
#include <cassert>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "dimensionvector.hh"

namespace morefit {

template <typename kernelT, typename evalT = double>
class EventVector {
  std::size_t nevents_ = 0;
  std::vector<dimension<evalT>*> dimensions_;
  unsigned int padding_ = 0;
  bool padded_ = false;
  int weight_idx_ = -1;
  std::vector<kernelT> data_;

  std::size_t column_length(std::size_t n, bool padded, unsigned int pad) const {
    return (padded && pad > 0) ? n + ((pad - (n % pad)) % pad) : n;
  }
  void relayout(std::size_t new_n, bool new_padded, unsigned int new_pad) {
    std::size_t old_len = column_length(nevents_, padded_, padding_);
    std::size_t new_len = column_length(new_n, new_padded, new_pad);
    if (old_len == new_len) return;
    std::vector<kernelT> nd(dimensions_.size() * new_len, kernelT(0));
    std::size_t keep = std::min(old_len, new_len);
    for (std::size_t d = 0; d < dimensions_.size(); d++)
      std::copy(data_.begin() + d * old_len, data_.begin() + d * old_len + keep, nd.begin() + d * new_len);
    data_.swap(nd);
  }

 public:
  EventVector() {}
  EventVector(std::vector<dimension<evalT>*> dimensions, unsigned long nevents = 0, bool padded = false, int padding = 0)
      : nevents_(nevents), dimensions_(dimensions), padding_(padding), padded_(padded) {
    data_.assign(dimensions_.size() * column_length(nevents_, padded_, padding_), kernelT(0));
  }
  int set_event_weight(std::string weight_name = "") {
    int idx = weight_name.empty() ? -1 : dim_index(weight_name);
    if (!weight_name.empty() && idx == -1) {
      std::cout << "Could not find event weight variable " << weight_name << " in vector." << std::endl;
      assert(0);
    }
    weight_idx_ = idx;
    return weight_idx_;
  }
  int event_weight_idx() const { return weight_idx_; }
  std::string event_weight_name() const { return weight_idx_ >= 0 ? dimensions_.at(weight_idx_)->get_name() : ""; }
  bool has_event_weight() const { return weight_idx_ >= 0; }
  std::size_t nevents_padded(std::size_t nevents) const { return column_length(nevents, true, padding_); }
  std::size_t nevents_padded() const { return column_length(nevents_, padded_, padding_); }
  unsigned int ndimensions() const { return (unsigned int)dimensions_.size(); }
  void set_padding(bool padded, unsigned int new_padding) {
    if (padded == padded_ && new_padding == padding_) return;
    relayout(nevents_, padded, new_padding);
    padded_ = padded;
    padding_ = new_padding;
  }
  bool is_padded() const { return padded_; }
  unsigned int padding() const { return padding_; }
  std::vector<dimension<evalT>*> get_dimensions() const { return dimensions_; }
  std::vector<dimension<evalT>> copy_dimensions() const {
    std::vector<dimension<evalT>> r;
    for (auto* d : dimensions_) r.push_back(*d);
    return r;
  }
  std::vector<std::string> get_dimensions_str() const {
    std::vector<std::string> r;
    for (auto* d : dimensions_) r.push_back(d->get_name());
    return r;
  }
  unsigned int size_t() const { return sizeof(kernelT); }
  const std::size_t& nevents() const { return nevents_; }
  void resize(std::size_t nevents) {
    if (nevents == nevents_) return;
    relayout(nevents, padded_, padding_);
    nevents_ = nevents;
  }
  void add_dimensions(std::vector<dimension<evalT>*> new_dimensions) {
    std::size_t len = nevents_padded();
    dimensions_.insert(dimensions_.end(), new_dimensions.begin(), new_dimensions.end());
    data_.resize(dimensions_.size() * len, kernelT(0));
  }
  std::size_t buffer_size() const { return nevents_padded() * ndimensions() * sizeof(kernelT); }
  int dim_index(std::string d) const {
    for (std::size_t i = 0; i < dimensions_.size(); i++)
      if (dimensions_[i]->get_name() == d) return (int)i;
    return -1;
  }
  kernelT* get_data() const { return const_cast<kernelT*>(data_.data()); }
  const kernelT& operator()(std::size_t n, unsigned int d) const { return data_[d * nevents_padded() + n]; }
  kernelT& operator()(std::size_t n, unsigned int d) { return data_[d * nevents_padded() + n]; }
  const kernelT& operator()(std::size_t n, std::string d) const {
    int idx = dim_index(d);
    assert(idx != -1);
    return data_[idx * nevents_padded() + n];
  }
  kernelT& operator()(std::size_t n, std::string d) {
    int idx = dim_index(d);
    assert(idx != -1);
    return data_[idx * nevents_padded() + n];
  }
  void print(int precision = 6, int width = 10, int max_events = -1, bool print_header = true, bool print_info = true, bool print_index = true) const {
    if (print_info)
      std::cout << "morefit::EventVector with " << ndimensions() << " dimensions and " << nevents() << " events, " << buffer_size() << " bytes" << std::endl;
    if (print_header) {
      if (print_index) std::cout << std::setw(width) << "index";
      for (auto* d : dimensions_) std::cout << std::setw(width) << d->get_name();
      std::cout << std::endl;
    }
    std::size_t m = max_events < 0 ? nevents_ : std::min<std::size_t>(nevents_, (std::size_t)max_events);
    for (std::size_t i = 0; i < m; i++) {
      if (print_index) std::cout << std::setw(width) << i;
      for (unsigned int j = 0; j < ndimensions(); j++) std::cout << std::setw(width) << std::setprecision(precision) << (*this)(i, j);
      std::cout << std::endl;
    }
    if (m < nevents_) std::cout << "... and " << nevents_ - m << " events more." << std::endl;
  }
};

}  // namespace morefit
#endif
