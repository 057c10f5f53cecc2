// Parameter files of the reference's drivers, read into the PARGEMSLR_IO_*
// parameter array.
//
// Two layouts, one value or keyword per line in a fixed line order:
//  - the sequential drivers' `inputs` (read_inputs_from_file,
//    ParGeMSLR/TESTS/sequential/io.cpp:336-700, 38 lines);
//  - the hypre interface's `inputs` (read_inputs_from_file,
//    ParGeMSLR/TESTS/hypre/pargemslr_hypre_interface.cpp:391-820, 31 lines,
//    some lines holding two to four numbers).
// Each line is described by a rule: a keyword table (the first
// space-delimited word, upper-cased, as PargemslrReadFirstWord,
// ParGeMSLR/SRC/utils/utils.cpp:336-358) or a list of numbers read from the
// start of the line as sscanf("%lf %lf ...") does (parsing stops at the first
// field that is not a number; earlier fields keep their new values, later ones
// their old values), followed by copies of GLOBAL entries into their LOCAL
// twins.  The copies reproduce the reference's, including its sequential
// reader setting LFIL_*_LOCAL from ROWNNZ_*_GLOBAL.  An unknown keyword stops
// the read with PGS_ERROR_IO and leaves the lines before it applied, as the
// reference does.  Lines past the last rule are ignored.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/pargemslr_amd.h"

namespace {

struct Keyword {
  const char* word;
  double value;
};

struct LineRule {
  std::vector<Keyword> words;                 // keyword line when not empty
  std::vector<int> dst;                       // receiving indices (-1: read and dropped)
  std::vector<std::pair<int, int>> copies;    // params[first] = params[second] afterwards
};

// option tables (enum orders of ParGeMSLR/SRC/preconditioners/gemslr.hpp:28-140, ilu.hpp:90-98)
const std::vector<Keyword> kBSolve = {{"ILUT", 0}, {"ILUK", 1}, {"GEMSLR", 2}};
const std::vector<Keyword> kPartition = {{"ND", 0}, {"RKWAY", 1}};
const std::vector<Keyword> kIluPerm = {{"RCM", 1}, {"AMD", 2}, {"ND", 3}, {"NO", 0}};
const std::vector<Keyword> kLowRank = {{"TR", 1}, {"STD", 0}, {"SUB", 2}};

std::vector<LineRule> sequential_rules() {
  auto num = [](int d, std::vector<std::pair<int, int>> c = {}) { return LineRule{{}, {d}, c}; };
  auto word = [](const std::vector<Keyword>& w, std::vector<int> d) { return LineRule{w, d, {}}; };
  return {
      word(kBSolve, {1}), num(2), word(kBSolve, {3}),
      num(11), num(12), num(13), num(14),
      num(20), num(21), num(22), num(23),
      word(kPartition, {24}), word(kPartition, {25}),
      num(40, {{41, 40}}), num(44, {{45, 44}}), num(48, {{49, 48}}), num(50, {{51, 50}}),
      num(52, {{53, 48}}), num(54, {{55, 50}}),
      word(kIluPerm, {58, 59}), word(kLowRank, {70, 72}), word(kLowRank, {71, 73}),
      num(74, {{76, 74}}), num(75, {{77, 75}}), num(86, {{88, 86}}), num(87, {{89, 87}}),
      num(78, {{80, 78}}), num(79, {{81, 79}}), num(82, {{84, 82}}), num(83, {{85, 83}}),
      num(90, {{92, 90}}), num(91, {{93, 91}}),
      num(130), num(131), num(132), num(142),
      word({{"LU", 0}, {"U", 1}, {"MUL", 2}}, {141}),
      num(150),
  };
}

std::vector<LineRule> hypre_rules() {
  auto nums = [](std::vector<int> d, std::vector<std::pair<int, int>> c = {}) { return LineRule{{}, d, c}; };
  auto word = [](const std::vector<Keyword>& w, std::vector<int> d) { return LineRule{w, d, {}}; };
  return {
      word({{"BJ", 0}, {"ESCHUR", 1}, {"GEMSLR", 2}, {"PSLR", 3}}, {0}),
      nums({34}), word(kBSolve, {1}), nums({2}), word(kBSolve, {3}),
      word({{"ILUT", 0}, {"ILUK", 1}, {"BJILUT", 2}, {"BJILUK", 3}}, {4}),
      nums({160}), nums({20, 21}), nums({22, 23}),
      word(kPartition, {24}), word(kPartition, {25}),
      nums({40, 42, 46, 44}, {{41, 40}, {43, 42}, {47, 46}, {45, 44}}),
      nums({48, 56, 50}, {{49, 48}, {57, 56}, {51, 50}}),
      nums({52, -1, 54}, {{53, 52}, {55, 54}}),
      word(kIluPerm, {58, 59}), word(kLowRank, {70, 72}), word(kLowRank, {71, 73}),
      nums({74, 75}, {{76, 74}, {77, 75}}), nums({86, 87}, {{88, 86}, {89, 87}}),
      nums({78, 79}, {{80, 78}, {81, 79}}), nums({82, 83}, {{84, 82}, {85, 83}}),
      nums({90, 91}, {{92, 90}, {93, 91}}),
      nums({130}), nums({131}), nums({132}), nums({142}),
      word({{"LU", 0}, {"U", 1}}, {141}),
      nums({143}), nums({144}),
      word({{"MILU", 0}, {"ILU", 1}}, {5}),
      nums({150}),
  };
}

std::string first_word(const std::string& line) {
  size_t i = 0;
  while (i < line.size() && line[i] == ' ') ++i;
  std::string w;
  while (i < line.size() && !std::isspace((unsigned char)line[i])) w += (char)std::toupper((unsigned char)line[i++]);
  return w;
}

bool apply_rule(const LineRule& r, const std::string& line, double* params) {
  if (!r.words.empty()) {
    const std::string w = first_word(line);
    for (const Keyword& k : r.words)
      if (w == k.word) {
        for (int d : r.dst) params[d] = k.value;
        return true;
      }
    return false;
  }
  const char* p = line.c_str();
  for (int d : r.dst) {
    char* e;
    const double v = std::strtod(p, &e);
    if (e == p) break;
    if (d >= 0) params[d] = v;
    p = e;
  }
  for (const auto& c : r.copies) params[c.first] = params[c.second];
  return true;
}

}  // namespace

extern "C" int pgs_params_read_inputs(const char* path, int format, double* params) {
  if (!path || !params || (format != PGS_INPUTS_SEQUENTIAL && format != PGS_INPUTS_HYPRE))
    return PGS_ERROR_INVALID_PARAM;
  FILE* f = std::fopen(path, "r");
  if (!f) return PGS_ERROR_IO;
  const std::vector<LineRule> rules = format == PGS_INPUTS_SEQUENTIAL ? sequential_rules() : hypre_rules();
  std::string line;
  size_t lineno = 0;
  int err = 0;
  for (;;) {
    line.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') line += (char)ch;
    if (ch == EOF && line.empty()) break;
    if (lineno < rules.size() && !apply_rule(rules[lineno], line, params)) {
      err = PGS_ERROR_IO;
      break;
    }
    ++lineno;
    if (ch == EOF) break;
  }
  std::fclose(f);
  return err;
}
