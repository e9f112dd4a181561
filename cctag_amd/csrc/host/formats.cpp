// The reference's wire and file formats on the drop-in boundary:
//  * the Parameters XML archive (boost::archive::xml_iarchive / oarchive of
//    cctag::Parameters::serialize, Params.hpp:210-248, read by
//    ICCTag.cpp:47-58 under the element name "CCTagsParams");
//  * the marker bank text file (CCTagMarkersBank.cpp:60-78 with the line
//    grammar of CCTagMarkersBank.hpp:42-65: per line, numbers written as an
//    unsigned ratio "p/q" or as a decimal, separated by white space).
// Host code, no device needed.
#include <cerrno>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/cctag_amd.h"

namespace cctagx {
void set_error(const std::string& msg);  // the calling thread's cctagx_last_error()
}

namespace {

using cctagx::set_error;

enum class Kind { F32, I32, Bool, Unused };

struct Field {
  const char* name;  // the serialized member name
  Kind kind;
  size_t offset;     // into cctagx_params (Unused: none)
};

#define CCTAGX_FIELD(n, k, m) {n, k, offsetof(cctagx_params, m)}
// Params.hpp:212-246, in archive order.
const Field kFields[] = {
    CCTAGX_FIELD("_cannyThrLow", Kind::F32, canny_thr_low),
    CCTAGX_FIELD("_cannyThrHigh", Kind::F32, canny_thr_high),
    CCTAGX_FIELD("_distSearch", Kind::I32, dist_search),
    CCTAGX_FIELD("_thrGradientMagInVote", Kind::I32, thr_gradient_mag_in_vote),
    CCTAGX_FIELD("_angleVoting", Kind::F32, angle_voting),
    CCTAGX_FIELD("_ratioVoting", Kind::F32, ratio_voting),
    CCTAGX_FIELD("_averageVoteMin", Kind::F32, average_vote_min),
    CCTAGX_FIELD("_thrMedianDistanceEllipse", Kind::F32, thr_median_distance_ellipse),
    CCTAGX_FIELD("_maximumNbSeeds", Kind::I32, maximum_nb_seeds),
    CCTAGX_FIELD("_maximumNbCandidatesLoopTwo", Kind::I32, maximum_nb_candidates_loop_two),
    CCTAGX_FIELD("_nCrowns", Kind::I32, n_crowns),
    CCTAGX_FIELD("_minPointsSegmentCandidate", Kind::I32, min_points_segment_candidate),
    CCTAGX_FIELD("_minVotesToSelectCandidate", Kind::I32, min_votes_to_select_candidate),
    CCTAGX_FIELD("_threshRobustEstimationOfOuterEllipse", Kind::F32,
                 thresh_robust_estimation_of_outer_ellipse),
    CCTAGX_FIELD("_ellipseGrowingEllipticHullWidth", Kind::F32, ellipse_growing_elliptic_hull_width),
    CCTAGX_FIELD("_windowSizeOnInnerEllipticSegment", Kind::I32,
                 window_size_on_inner_elliptic_segment),
    CCTAGX_FIELD("_numberOfMultiresLayers", Kind::I32, number_of_multires_layers),
    CCTAGX_FIELD("_numberOfProcessedMultiresLayers", Kind::I32, number_of_processed_multires_layers),
    CCTAGX_FIELD("_nSamplesOuterEllipse", Kind::I32, n_samples_outer_ellipse),
    CCTAGX_FIELD("_numCutsInIdentStep", Kind::I32, num_cuts_in_ident_step),
    CCTAGX_FIELD("_numSamplesOuterEdgePointsRefinement", Kind::I32,
                 num_samples_outer_edge_points_refinement),
    {"_cutsSelectionTrials", Kind::Unused, 0},
    CCTAGX_FIELD("_sampleCutLength", Kind::I32, sample_cut_length),
    CCTAGX_FIELD("_imagedCenterNGridSample", Kind::I32, imaged_center_n_grid_sample),
    CCTAGX_FIELD("_imagedCenterNeighbourSize", Kind::F32, imaged_center_neighbour_size),
    CCTAGX_FIELD("_minIdentProba", Kind::F32, min_ident_proba),
    {"_useLMDif", Kind::Unused, 0},
    CCTAGX_FIELD("_searchForAnotherSegment", Kind::Bool, search_for_another_segment),
    {"_writeOutput", Kind::Unused, 0},
    CCTAGX_FIELD("_doIdentification", Kind::Bool, do_identification),
    {"_maxEdges", Kind::Unused, 0},
    {"_useCuda", Kind::Unused, 0},
    {"_pinnedCounters", Kind::Unused, 0},
    {"_pinnedNearbyPoints", Kind::Unused, 0},
};
#undef CCTAGX_FIELD

// Values the archive holds for the members the C ABI has no field for
// (Params.hpp:27-66 defaults).
const char* unused_default(const char* name) {
  if (!std::strcmp(name, "_cutsSelectionTrials")) return "500";
  if (!std::strcmp(name, "_useLMDif")) return "1";
  if (!std::strcmp(name, "_writeOutput")) return "0";
  if (!std::strcmp(name, "_maxEdges")) return "20000";
  if (!std::strcmp(name, "_useCuda")) return "1";
  if (!std::strcmp(name, "_pinnedCounters")) return "100";
  if (!std::strcmp(name, "_pinnedNearbyPoints")) return "7000";
  return "0";
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// The text of element <name ...>text</name> at or after `from` inside
// [from, end); npos when absent.
size_t element(const std::string& x, const std::string& name, size_t from, size_t end,
               std::string& text, size_t* after = nullptr) {
  const std::string open = "<" + name;
  for (size_t p = x.find(open, from); p != std::string::npos && p < end;
       p = x.find(open, p + 1)) {
    const char c = x[p + open.size()];
    if (c != '>' && c != ' ' && c != '\t' && c != '\n' && c != '\r' && c != '/') continue;
    const size_t gt = x.find('>', p);
    if (gt == std::string::npos || gt >= end) return std::string::npos;
    if (x[gt - 1] == '/') {  // <name/>
      text.clear();
      if (after) *after = gt + 1;
      return p;
    }
    const std::string close = "</" + name + ">";
    const size_t q = x.find(close, gt + 1);
    if (q == std::string::npos || q > end) return std::string::npos;
    text = x.substr(gt + 1, q - gt - 1);
    if (after) *after = q + close.size();
    return p;
  }
  return std::string::npos;
}

bool parse_number(const std::string& t, Kind k, double& v) {
  const std::string s = trim(t);
  if (s.empty()) return false;
  char* e = nullptr;
  errno = 0;
  if (k == Kind::F32) v = std::strtod(s.c_str(), &e);
  else v = (double)std::strtoll(s.c_str(), &e, 10);
  return e && *e == 0 && errno == 0;
}

int parse_params(const std::string& x, cctagx_params* out) {
  std::string body;
  size_t end_of_root = 0;
  if (element(x, "CCTagsParams", 0, x.size(), body, &end_of_root) == std::string::npos) {
    set_error("no CCTagsParams element in the parameter archive");
    return CCTAGX_ERR_ARG;
  }
  cctagx_params p;
  cctagx_params_default(&p, 3);
  for (const Field& f : kFields) {
    std::string t;
    if (element(body, f.name, 0, body.size(), t) == std::string::npos) {
      set_error(std::string("parameter archive lacks ") + f.name);
      return CCTAGX_ERR_ARG;
    }
    double v = 0;
    if (!parse_number(t, f.kind == Kind::F32 ? Kind::F32 : Kind::I32, v)) {
      set_error(std::string("parameter ") + f.name + " is not a number: " + trim(t));
      return CCTAGX_ERR_ARG;
    }
    char* base = reinterpret_cast<char*>(&p);
    switch (f.kind) {
      case Kind::F32: *reinterpret_cast<float*>(base + f.offset) = (float)v; break;
      case Kind::I32:
      case Kind::Bool: *reinterpret_cast<int32_t*>(base + f.offset) = (int32_t)v; break;
      default: break;
    }
  }
  *out = p;
  return CCTAGX_OK;
}

// A float with the digits that read back to the same value.
std::string float_text(float v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.9g", (double)v);
  return buf;
}

std::string params_xml(const cctagx_params& p) {
  std::ostringstream o;
  o << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"yes\" ?>\n"
    << "<!DOCTYPE boost_serialization>\n"
    << "<boost_serialization signature=\"serialization::archive\" version=\"17\">\n"
    << "<CCTagsParams class_id=\"0\" tracking_level=\"0\" version=\"0\">\n";
  const char* base = reinterpret_cast<const char*>(&p);
  for (const Field& f : kFields) {
    o << "\t<" << f.name << ">";
    switch (f.kind) {
      case Kind::F32: o << float_text(*reinterpret_cast<const float*>(base + f.offset)); break;
      case Kind::I32: o << *reinterpret_cast<const int32_t*>(base + f.offset); break;
      case Kind::Bool: o << (*reinterpret_cast<const int32_t*>(base + f.offset) != 0 ? 1 : 0); break;
      default: o << unused_default(f.name); break;
    }
    o << "</" << f.name << ">\n";
  }
  o << "</CCTagsParams>\n</boost_serialization>\n\n";
  return o.str();
}

// CCTagMarkersBank::cctagLineParse: a sequence of `uint '/' uint` (their
// float ratio) or `double` items separated by white space; parsing stops at
// the first text that is neither, keeping what was read before it.
void parse_bank_line(const char* s, std::vector<float>& rr) {
  const char* p = s;
  for (;;) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
    if (!*p) return;
    // uint '/' uint
    if (*p >= '0' && *p <= '9') {
      const char* q = p;
      unsigned long a = 0;
      while (*q >= '0' && *q <= '9') a = a * 10 + (unsigned long)(*q++ - '0');
      const char* r = q;
      while (*r == ' ' || *r == '\t') ++r;
      if (*r == '/') {
        ++r;
        while (*r == ' ' || *r == '\t') ++r;
        if (*r >= '0' && *r <= '9') {
          unsigned long b = 0;
          while (*r >= '0' && *r <= '9') b = b * 10 + (unsigned long)(*r++ - '0');
          float n = (float)a;
          n = n / (float)b;
          rr.push_back(n);
          p = r;
          continue;
        }
      }
    }
    // double
    char* e = nullptr;
    const double d = std::strtod(p, &e);
    if (e == p) return;
    rr.push_back((float)d);
    p = e;
  }
}

int parse_bank(const std::string& text, float* ratios, int max_rows, int max_cols, int* rows,
               int* cols) {
  std::vector<std::vector<float>> bank;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    std::vector<float> rr;
    parse_bank_line(line.c_str(), rr);
    if (!rr.empty()) bank.push_back(rr);
  }
  if (bank.empty()) {
    set_error("the bank holds no row");
    return CCTAGX_ERR_ARG;
  }
  for (const auto& r : bank)
    if (r.size() != bank[0].size()) {
      set_error("bank rows of different lengths are not supported");
      return CCTAGX_ERR_UNSUPPORTED;
    }
  *rows = (int)bank.size();
  *cols = (int)bank[0].size();
  if (!ratios) return CCTAGX_OK;  // size query
  if (*rows > max_rows || *cols > max_cols) return CCTAGX_ERR_CAPACITY;
  for (int i = 0; i < *rows; ++i)
    for (int j = 0; j < *cols; ++j) ratios[(size_t)i * *cols + j] = bank[i][j];
  return CCTAGX_OK;
}

bool read_file(const char* path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) return false;
  std::ostringstream s;
  s << f.rdbuf();
  out = s.str();
  return true;
}

}  // namespace

extern "C" {

int cctagx_params_parse_xml(const char* text, cctagx_params* out) {
  if (!text || !out) return CCTAGX_ERR_ARG;
  return parse_params(text, out);
}

int cctagx_params_read_xml(const char* path, cctagx_params* out) {
  if (!path || !out) return CCTAGX_ERR_ARG;
  std::string x;
  if (!read_file(path, x)) {
    set_error(std::string("The input parameter file \"") + path + "\" is missing");
    return CCTAGX_ERR_ARG;
  }
  return parse_params(x, out);
}

int cctagx_params_write_xml(const cctagx_params* p, char* buf, size_t cap, size_t* len) {
  if (!p || !len) return CCTAGX_ERR_ARG;
  const std::string x = params_xml(*p);
  *len = x.size();
  if (!buf) return CCTAGX_OK;
  if (cap < x.size() + 1) return CCTAGX_ERR_CAPACITY;
  std::memcpy(buf, x.c_str(), x.size() + 1);
  return CCTAGX_OK;
}

int cctagx_bank_parse(const char* text, float* ratios, int max_rows, int max_cols, int* rows,
                      int* cols) {
  if (!text || !rows || !cols) return CCTAGX_ERR_ARG;
  return parse_bank(text, ratios, max_rows, max_cols, rows, cols);
}

int cctagx_bank_read(const char* path, float* ratios, int max_rows, int max_cols, int* rows,
                     int* cols) {
  if (!path || !rows || !cols) return CCTAGX_ERR_ARG;
  std::string x;
  if (!read_file(path, x)) {
    set_error(std::string("Unable to open the bank file: ") + path);
    return CCTAGX_ERR_ARG;
  }
  return parse_bank(x, ratios, max_rows, max_cols, rows, cols);
}

}  // extern "C"
