/**
 * @file plotter.hh
 * One-dimensional projections of a density over binned data, behind the reference's plotter
 * interface (plotter.hh:44-216 for plotter_options and make_kernels, 218-640 for plot):
 * plotter_options and plotter<kernelT, evalT, backendT, computeT>.
 *
 *  integral kernel  the density's integral over every fine bin of the plotted observable (the
 *                   other observables over their full ranges, every parameter at its current
 *                   value) from PDF::definite_integral, as one store kernel "integral_kernel"
 *                   over a two-column input of bin edges (morefit_from, morefit_to) on the
 *                   backend's block: on the MI355X one work item per fine bin, in the CPU
 *                   backend one loop iteration per fine bin.
 *  host work        the data histogram (the data EventVector is a host container), the scale
 *                   of the density to the histogram, residuals and pulls, and the output file:
 *                   a matplotlib script (.py) or a ROOT macro (.C), written as the reference
 *                   writes them, including its half-length data error bars in the matplotlib
 *                   script (plotter.hh:381-383); TeX labels are raw string literals there.
 *                   A ROOT canvas (.eps, .pdf through ROOT) needs ROOT, which this build does
 *                   not link: such a file is reported as unsupported, as the reference
 *                   without WITH_ROOT does.
 *
 * The projection of the last plot() stays in `last` (fine-bin integrals, counts, errors,
 * residuals, pulls).  For inputs the reference cannot handle: an event at exactly range_max
 * counts in the last bin and events outside [range_min, range_max] are not counted (the
 * reference indexes past its histogram for both).
 */
#ifndef MOREFIT_PLOTTER_HH
#define MOREFIT_PLOTTER_HH

#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "compute.hh"
#include "eventvector.hh"
#include "graph.hh"
#include "parametervector.hh"
#include "pdf.hh"

namespace morefit {

struct plotter_options {
  enum plotter_type { Root, MatPlotLib };
  plotter_type plotter = Root;
  int pdf_bin_multiplier = 10;
  std::string file_format = "pdf";
  bool optimize_parameters = false;
  float buffering_cost_threshold = 2.0f;
  int print_level = 2;
  int plot_width = 1200;
  float aspect_ratio = 4.0f / 3.0f;
  bool plot_pulls = false;
  float pull_fraction = 0.20f;
  float max_pull = 3.0f;
  void print() const {
    const int w = 40;
    std::cout << "PLOTTER OPTIONS\n"
              << std::setw(w) << std::left << "  plotter type " << (plotter == Root ? "Root" : "MatPlotLib") << "\n"
              << std::setw(w) << std::left << "  pdf bin multiplier " << pdf_bin_multiplier << "\n"
              << std::setw(w) << std::left << "  file format " << file_format << "\n"
              << std::setw(w) << std::left << "  optimize parameters " << (optimize_parameters ? "YES" : "NO") << "\n"
              << std::setw(w) << std::left << "  buffering cost threshold " << buffering_cost_threshold << "\n"
              << std::setw(w) << std::left << "  print level " << print_level << "\n"
              << std::setw(w) << std::left << "  plot width " << plot_width << "\n"
              << std::setw(w) << std::left << "  aspect ratio " << aspect_ratio << "\n"
              << std::setw(w) << std::left << "  plot pulls " << (plot_pulls ? "YES" : "NO") << "\n"
              << std::setw(w) << std::left << "  pull fraction " << pull_fraction << "\n"
              << std::setw(w) << std::left << "  maximum pull range " << max_pull << std::endl;
  }
};

template <typename evalT>
struct projection {
  unsigned int nbins = 0, npdf_bins = 0;
  evalT range_min = 0, range_max = 0;
  std::vector<evalT> pdf_integrals;  // per fine bin: integral of the normalised density
  std::vector<evalT> counts, errors;  // per bin: (weighted) events and their error
  std::vector<evalT> residuals, pulls;
  double pdf_scale = 0.0;  // fine-bin integral -> events per bin
  double max_pull = 0.0;   // range of the pull panel
};

template <typename kernelT, typename evalT, typename backendT, typename computeT>
class plotter {
  typedef GraphPtr<kernelT, evalT> Ptr;
  plotter_options* opts_;
  backendT* backend_;
  computeT block_;
  dimension<evalT> res_dim_, from_dim_, to_dim_;
  EventVector<kernelT, evalT> res_buffer_, ranges_buffer_;
  std::vector<Ptr> buffer_expressions_;
  std::vector<std::string> buffer_names_;
  std::string variable_;

  // '#' of ROOT's TeX labels -> '\' for matplotlib's TeX
  static std::string tex_for_matplotlib(const std::string& s) {
    std::string r;
    for (char c : s) r += (c == '#') ? '\\' : c;
    return r;
  }
  template <typename V, typename F>
  static void write_list(std::ostream& os, const V& v, F f) {
    for (size_t i = 0; i < v.size(); i++) os << f(v[i]) << (i + 1 < v.size() ? ", " : "");
  }

 public:
  projection<evalT> last;
  double last_ms = 0.0, kernel_ms = 0.0;

  plotter(plotter_options* options, backendT* backend)
      : opts_(options),
        backend_(backend),
        block_(backend),
        res_dim_("integral", 0.0, 1.0),
        from_dim_("morefit_from", 0.0, 1.0),
        to_dim_("morefit_to", 0.0, 1.0),
        res_buffer_({&res_dim_}),
        ranges_buffer_({&from_dim_, &to_dim_}) {}

  // the integral kernel: definite_integral with every parameter at its value, the other
  // observables' boundaries at their ranges and the plotted observable's boundaries renamed
  // to the input columns morefit_from / morefit_to (reference plotter.hh:138-216)
  bool make_kernels(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<kernelT, evalT>* data) {
    std::vector<std::string> floating, names, dnames;
    std::vector<double> values, dvalues;
    for (auto* p : params) {
      if (!p->is_constant()) floating.push_back(p->get_name());
      names.push_back(p->get_name());
      values.push_back(p->get_value());
    }
    std::string vfrom = variable_, vto = variable_;
    for (auto* d : pdf->dimensions()) {
      if (d->get_name() == variable_) {
        vfrom = d->get_from_name();
        vto = d->get_to_name();
        continue;
      }
      dnames.push_back(d->get_from_name());
      dvalues.push_back(d->get_min());
      dnames.push_back(d->get_to_name());
      dvalues.push_back(d->get_max());
    }
    Ptr integral = pdf->definite_integral()->substitute(names, values)->simplify()->substitute(dnames, dvalues)->simplify();
    integral->rename_variable(vfrom, from_dim_.get_name());
    integral->rename_variable(vto, to_dim_.get_name());
    buffer_expressions_.clear();
    buffer_names_.clear();
    std::vector<Ptr> graphs;
    if (opts_->optimize_parameters) {
      graphs.emplace_back(integral->optimize_buffering_constant_terms(buffer_names_, buffer_expressions_, floating, "morefit_parambuffer_",
                                                                      opts_->buffering_cost_threshold));
      if (opts_->print_level > 1) {
        std::cout << "GRAPH: " << graphs.back()->get_kernel() << "\n";
        for (size_t i = 0; i < buffer_expressions_.size(); i++)
          std::cout << "BUFFER " << buffer_names_[i] << " = " << buffer_expressions_[i]->get_kernel() << "\n";
      }
    } else {
      graphs.emplace_back(integral->copy());
    }
    std::vector<std::string> pnames(floating);
    pnames.insert(pnames.end(), buffer_names_.begin(), buffer_names_.end());
    block_.SetupInputBuffer(ranges_buffer_.buffer_size());
    block_.SetupOutputBuffer(res_buffer_.buffer_size());
    block_.SetupParameterBuffer(std::max<size_t>(1, pnames.size()) * sizeof(kernelT));
    const auto t0 = std::chrono::steady_clock::now();
    block_.MakeComputeKernel("integral_kernel", (unsigned int)ranges_buffer_.nevents(), ranges_buffer_.copy_dimensions(),
                             res_buffer_.copy_dimensions(), pnames, graphs, false);
    block_.Finish();
    kernel_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (opts_->print_level > 1) std::cout << "kernel (integral) takes " << kernel_ms << " ms in total" << std::endl;
    return true;
  }

  // projection of `pdf` onto `variable` with the histogram of `data`: nbins bins on
  // [range_min, range_max] (the observable's range when both are equal), the density in
  // nbins * pdf_bin_multiplier fine bins; replot reuses the kernel of the previous call
  bool plot(PDF<kernelT, evalT>* pdf, std::vector<parameter<evalT>*> params, EventVector<kernelT, evalT>* data, std::string variable,
            std::string filename, std::string output_type, unsigned int nbins = 100, evalT range_min = 0.0, evalT range_max = 0.0,
            bool replot = false) {
    const auto t0 = std::chrono::steady_clock::now();
    variable_ = variable;
    dimension<evalT>* dim = nullptr;
    int varidx = -1;
    const auto& dims = pdf->dimensions();
    for (size_t i = 0; i < dims.size(); i++)
      if (dims[i]->get_name() == variable) {
        dim = dims[i];
        varidx = (int)i;
        break;
      }
    if (!dim) throw std::invalid_argument("plotter: variable " + variable + " is not an observable of the density");
    if (nbins == 0) throw std::invalid_argument("plotter: no bins");
    if (opts_->pdf_bin_multiplier < 1) throw std::invalid_argument("plotter: pdf_bin_multiplier must be at least 1");
    if (range_min == range_max) {
      range_min = dim->get_min();
      range_max = dim->get_max();
    }
    const unsigned int mult = (unsigned int)opts_->pdf_bin_multiplier, nfine = nbins * mult;
    const int pad = backend_->required_padding();
    ranges_buffer_.set_padding(pad != 0, pad);
    res_buffer_.set_padding(pad != 0, pad);
    ranges_buffer_.resize(nfine);
    res_buffer_.resize(nfine);
    const evalT dx = (range_max - range_min) / evalT(nfine);
    for (unsigned int i = 0; i < nfine; i++) {
      ranges_buffer_(i, 0u) = (kernelT)(range_min + i * dx);
      ranges_buffer_(i, 1u) = (kernelT)(range_min + (i + 1) * dx);
    }
    if (!replot) make_kernels(pdf, params, data);
    block_.SetupInputBuffer(ranges_buffer_.buffer_size());
    block_.SetupOutputBuffer(res_buffer_.buffer_size());
    block_.SetNevents((int)ranges_buffer_.nevents(), (int)ranges_buffer_.nevents_padded());
    block_.CopyToInputBuffer(ranges_buffer_);
    std::vector<std::string> fnames;
    std::vector<double> fvalues;
    std::vector<kernelT> pbuf;
    for (auto* p : params)
      if (!p->is_constant()) {
        fnames.push_back(p->get_name());
        fvalues.push_back(p->get_value());
        pbuf.push_back((kernelT)p->get_value());
      }
    for (auto& e : buffer_expressions_) pbuf.push_back((kernelT)e->eval(fnames, fvalues));
    if (pbuf.empty()) pbuf.push_back(kernelT(0));
    block_.CopyToParameterBuffer(pbuf);
    block_.SubmitKernel();
    block_.Finish();
    block_.CopyFromOutputBuffer(res_buffer_);

    projection<evalT>& P = last;
    P = projection<evalT>();
    P.nbins = nbins;
    P.npdf_bins = nfine;
    P.range_min = range_min;
    P.range_max = range_max;
    P.pdf_integrals.resize(nfine);
    for (unsigned int i = 0; i < nfine; i++) P.pdf_integrals[i] = (evalT)res_buffer_(i, 0u);
    P.counts.assign(nbins, 0.0);
    P.errors.assign(nbins, 0.0);
    const bool weighted = data->has_event_weight();
    const int widx = data->event_weight_idx();
    for (size_t i = 0; i < data->nevents(); i++) {
      const double x = (double)(*data)(i, (unsigned)varidx);
      if (!(x >= range_min && x <= range_max)) continue;
      long b = (long)std::floor((x - range_min) / (range_max - range_min) * nbins);
      if (b >= (long)nbins) b = (long)nbins - 1;
      const double w = weighted ? (double)(*data)(i, (unsigned)widx) : 1.0;
      P.counts[b] += w;
      P.errors[b] += w * w;
    }
    for (auto& e : P.errors) e = std::sqrt(e);
    double hist_int = 0.0, pdf_int = 0.0;
    for (auto c : P.counts) hist_int += c;
    for (auto v : P.pdf_integrals) pdf_int += v;
    P.pdf_scale = hist_int / pdf_int * double(nfine) / double(nbins);
    P.residuals.assign(nbins, 0.0);
    P.pulls.assign(nbins, 0.0);
    for (unsigned int i = 0; i < nfine; i++) {
      const unsigned b = i / mult;
      P.residuals[b] += (P.counts[b] - P.pdf_integrals[i] * P.pdf_scale) / mult;
    }
    for (unsigned int b = 0; b < nbins; b++) P.pulls[b] = P.residuals[b] / P.errors[b];
    if (opts_->max_pull > 0.0f) {
      P.max_pull = opts_->max_pull;
    } else {
      for (auto p : P.pulls) P.max_pull = std::max(P.max_pull, std::fabs((double)p));
      P.max_pull *= 1.05;
    }
    if (!filename.empty()) write(filename, output_type, *dim);
    last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (opts_->print_level > 1) std::cout << "plotting of " << data->nevents() << " events took " << last_ms << " ms in total" << std::endl;
    return true;
  }

 private:
  void write(const std::string& filename, const std::string& output_type, const dimension<evalT>& dim) const {
    const projection<evalT>& P = last;
    const std::filesystem::path path(filename);
    const std::string ext = path.extension().string(), stem = path.stem().string();
    const double lo = P.range_min, hi = P.range_max, w = (hi - lo) / P.nbins, wf = (hi - lo) / P.npdf_bins;
    double hist_max = 0.0;
    for (auto c : P.counts) hist_max = std::max(hist_max, (double)c);
    if (opts_->plotter == plotter_options::MatPlotLib && ext == ".py") {
      std::ofstream os(filename);
      const std::string tex = tex_for_matplotlib(dim.get_tex());
      os << "import matplotlib.pyplot as plt\nimport numpy as np\nplt.rcParams['text.usetex'] = True\n";
      os << "x = np.linspace(" << lo + 0.5 * w << ", " << hi - 0.5 * w << ", " << P.nbins << ")\n";
      os << "dx = [" << 0.5 * w << " for x in range(" << P.nbins << ")]\n";
      os << "y = [";
      write_list(os, P.counts, [](evalT v) { return v; });
      os << "]\ndy = [";
      write_list(os, P.errors, [](evalT v) { return 0.5 * v; });
      os << "]\n";
      if (opts_->plot_pulls) {
        os << "y_pull = [";
        write_list(os, P.pulls, [](evalT v) { return v; });
        os << "]\n";
      }
      os << "x_pdf = np.linspace(" << lo + 0.5 * wf << ", " << hi - 0.5 * wf << ", " << P.npdf_bins << ")\n";
      os << "y_pdf = [";
      const double sc = P.pdf_scale;
      write_list(os, P.pdf_integrals, [sc](evalT v) { return v * sc; });
      os << "]\n";
      const std::string ax = opts_->plot_pulls ? "ax[0]" : "ax";
      if (opts_->plot_pulls)
        os << "fig, ax = plt.subplots(nrows=2, ncols=1,  layout='constrained', height_ratios=[" << 1.0 - opts_->pull_fraction << ", "
           << opts_->pull_fraction << "])\n";
      else
        os << "fig, ax = plt.subplots()\n";
      os << ax << ".plot(x_pdf, y_pdf, linestyle='solid', color='red')\n";
      os << ax << ".errorbar(x, y, xerr=dx, yerr=dy, linestyle='none', marker='o', markersize=2, color='black')\n";
      os << ax << ".set_xlabel(r'$" << tex << "$')\n";
      os << ax << ".set_ylabel(r'$\\mathrm{events}$')\n";
      if (opts_->plot_pulls) {
        os << "ax[0].set_xlim(" << lo << ", " << hi << ")\n";
        os << "ax[0].set_ylim(" << 0.0 << ", " << hist_max * 1.15 << ")\n";
        os << "ax[1].bar(x=x, height=y_pull, width=" << w << ", facecolor='gray', linestyle='solid', edgecolor='black')\n";
        os << "#ax[1].set_xlabel(r'$" << tex << "$')\n";
        os << "ax[1].set_ylabel(r'$\\mathrm{pull}$')\n";
        os << "ax[1].set_xlim(" << lo << ", " << hi << ")\n";
        os << "ax[1].set_ylim(-" << P.max_pull << ", " << P.max_pull << ")\n";
      } else {
        os << "plt.xlim(" << lo << ", " << hi << ")\n";
        os << "plt.ylim(" << 0.0 << ", " << hist_max * 1.15 << ")\n";
      }
      os << "plt.show()\n";
      os << "fig.savefig(\"" << stem << ".pdf\", bbox_inches='tight')\n";
      if (opts_->print_level > 1) std::cout << "Written plot of variable " << dim.get_name() << " to file " << filename << std::endl;
    } else if (opts_->plotter == plotter_options::Root && ext == ".C") {
      std::ofstream os(filename);
      const std::string tex = dim.get_tex();
      auto array = [&os](const char* name, const std::vector<evalT>& v) {
        os << "double " << name << "[" << v.size() << "] = {";
        write_list(os, v, [](evalT x) { return x; });
        os << "};\n";
      };
      auto fill = [&os](const char* hist, const char* values, const char* errors, size_t n) {
        os << "for (unsigned int i=0; i<" << n << "; i++)\n{\n  " << hist << "->SetBinContent(i+1, " << values << "[i]);\n  " << hist
           << "->SetBinError(i+1, " << errors << ");\n}\n";
      };
      os << "int " << stem << "() {\n";
      os << "gROOT->SetStyle(\"Plain\");\ngStyle->SetOptFit(0);\ngStyle->SetOptStat(0);\ngStyle->SetTextFont(132);\n"
         << "gStyle->SetTextSize(0.06);\ngStyle->SetTitleFont(132,\"xyz\");\ngStyle->SetLabelFont(132,\"xyz\");\n"
         << "gStyle->SetLabelSize(0.05,\"xyz\");\ngStyle->SetTitleSize(0.06,\"xyz\");\n";
      os << "TCanvas* c0 = new TCanvas(\"c0\", \"c0\", " << opts_->plot_width << ", " << opts_->plot_width / evalT(opts_->aspect_ratio) << ");\n";
      os << "TH1D* hist = new TH1D(\"hist\", \";" << tex << ";events\", " << P.nbins << ", " << lo << ", " << hi << ");\n";
      array("hist_values", P.counts);
      array("hist_errors", P.errors);
      fill("hist", "hist_values", "hist_errors[i]", P.nbins);
      os << "TH1D* hist_pdf = new TH1D(\"hist_pdf\", \";" << tex << ";events\", " << P.npdf_bins << ", " << lo << ", " << hi << ");\n";
      array("pdf_values", P.pdf_integrals);
      array("pdf_errors", std::vector<evalT>(P.npdf_bins, 0.0));
      fill("hist_pdf", "pdf_values", "pdf_errors[i]", P.npdf_bins);
      os << "hist_pdf->Scale(hist->Integral()/hist_pdf->Integral()*" << P.npdf_bins << "/double(" << P.nbins << "));\n";
      const double pf = opts_->pull_fraction;
      if (opts_->plot_pulls) {
        os << "TH1D* pull_hist = new TH1D(\"pull_hist\", \";" << tex << ";pull\", " << P.nbins << ", " << lo << ", " << hi << ");\n";
        array("pull_values", P.pulls);
        fill("pull_hist", "pull_values", "0.0", P.nbins);
        os << "c0->cd(0);\nTPad *hist_pad = new TPad(\"hist_pad\",\"\",0.0, " << pf << ", 1.0, 1.0);\n"
           << "hist_pad->SetMargin(0.125,0.05,0.125,0.05);\nhist_pad->Draw();\nhist_pad->cd();\n"
           << "hist->SetMinimum(0.0);\nhist->Draw(\"e\");\nhist_pdf->SetLineColor(2);\nhist_pdf->SetLineWidth(2.0);\n"
           << "hist_pdf->Draw(\"lcsame\");\nc0->cd(0);\n"
           << "TPad *pull_pad = new TPad(\"pull_pad\",\"\",0.0, 0.0, 1.0, " << pf << ");\n"
           << "pull_pad->SetMargin(0.125,0.05,0.125,0.125);\npull_pad->Draw();\npull_pad->cd();\n"
           << "pull_hist->SetFillColor(kGray);\n"
           << "pull_hist->GetYaxis()->SetTitleOffset(1.0*" << pf / (1.0 - pf) << ");\n"
           << "pull_hist->GetYaxis()->SetLabelSize(0.05*" << (1.0 - pf) / pf << ");\n"
           << "pull_hist->GetYaxis()->SetTitleSize(0.06*" << (1.0 - pf) / pf << ");\n"
           << "pull_hist->GetXaxis()->SetLabelSize(0);\npull_hist->GetXaxis()->SetTitleSize(0);\n"
           << "pull_hist->SetMinimum(-" << P.max_pull << ");\npull_hist->SetMaximum(+" << P.max_pull << ");\n"
           << "pull_hist->Draw(\"hist\");\n"
           << "TGaxis *axis = new TGaxis(pull_hist->GetXaxis()->GetXmin(), pull_hist->GetMaximum(), pull_hist->GetXaxis()->GetXmax(), "
              "pull_hist->GetMaximum(), pull_hist->GetXaxis()->GetXmin(), pull_hist->GetXaxis()->GetXmax(), 510, \"-U\");\n"
           << "axis->Draw();\nc0->Print(\"" << stem << ".eps\", \"eps\");\n"
           << "delete hist;\ndelete hist_pdf;\ndelete hist_pad;\ndelete pull_hist;\ndelete pull_pad;\ndelete c0;\n";
      } else {
        os << "c0->cd()->SetMargin(0.125,0.05,0.125,0.05);\nhist->SetMinimum(0.0);\nhist->Draw(\"e\");\n"
           << "hist_pdf->SetLineColor(2);\nhist_pdf->SetLineWidth(2.0);\nhist_pdf->Draw(\"lcsame\");\n"
           << "c0->Print(\"" << stem << ".eps\", \"eps\");\ndelete hist;\ndelete hist_pdf;\ndelete c0;\n";
      }
      os << "return 1;\n}\n";
      if (opts_->print_level > 1) std::cout << "Written plot of variable " << dim.get_name() << " to file " << filename << std::endl;
    } else {
      std::cout << "Unknown file extension: " << filename << (opts_->plotter == plotter_options::Root && !output_type.empty()
                                                                   ? " (ROOT canvases need a ROOT build)"
                                                                   : "")
                << std::endl;
    }
  }
};

}  // namespace morefit
#endif
