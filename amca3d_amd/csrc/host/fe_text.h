// FE temperature frames in the reference's text format (README "PBF_x";
// FiniteElementManager::load_mesh_and_first_two_steps_results_from_txt_file and
// load_next_step_nodal_data, LoadDataFromFile.cpp:21-191).  A file holds a
// title block, then frames: a subtitle block whose fourth line reads
// "step step time time x0 x1 y0 y1 z0 z1", and one "x y z T" line per node of
// a structured hex mesh (x fastest).  Frames after the first start with an
// empty line.
#pragma once
#include <string>
#include <vector>

namespace amca_host {

struct FeTextOptions {
  int title_lines = 0;        // lines_for_title
  int subtitle_lines = 5;     // lines_for_subtitle (the 4th holds step/time/range)
  double length_scale = 1.0;  // coordinates: (x + offset) * length_scale
  double time_scale = 1.0;    // times: (t - t_first) * time_scale, first frame at 0
  double offset[3] = {0, 0, 0};
};

struct FeFrames {
  int fnx = 0, fny = 0, fnz = 0;
  std::vector<double> xyz;     // 3 per node (after offset and scale)
  std::vector<double> times;   // per frame, relative (frame 0 at 0)
  std::vector<double> theta;   // [frame][node]
  int nodes() const { return fnx * fny * fnz; }
  int frames() const { return (int)times.size(); }
};

// Reads every frame (load_data_from_file for_whole_time); throws
// std::runtime_error with the reference's messages on malformed files.
FeFrames load_fe_text(const std::string& path, const FeTextOptions& o);

// load_fe_text through a binary cache beside the file (<path>.amcacache): a
// cache whose source size, modification time and parse options match is read
// instead of the text; otherwise the text is parsed and the cache (re)written
// (a cache that cannot be written is skipped).  The frames are the same.
FeFrames load_fe_frames_cached(const std::string& path, const FeTextOptions& o);

// Writes frames in the same format (raw coordinates and times are written as
// given; `times` absolute).  Title: `title_lines` lines.
void write_fe_text(const std::string& path, int title_lines, int fnx, int fny, int fnz,
                   const double* xyz, int n_frames, const double* times, const double* theta);

}  // namespace amca_host
