// Drop-in svo_builder / svo_builder_binary (SVO_PAYLOAD selects the flavour, as the
// reference's BINARY_VOXELIZATION define does, CMakeLists.txt:39-49).
//
// Same flags, messages and exit statuses as main.cpp:100-196,244-399; same .tri/.tridata
// input (tri_tools.h:75-115) and .octree/.octreenodes/.octreedata output
// (octree_io.h:49-83) with output base <tri base><gridsize>_<partitions>
// (partitioner.cpp:88-91,141). The stages run on the GPU through the C ABI; the mesh is
// read once and stays in HBM, so no .trip/.tripdata files are written (the reference
// deletes them before exiting, partitioner.cpp:31-40).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "svo_amd.h"

using std::cout;
using std::endl;
using std::string;

#ifndef SVO_PAYLOAD
#define SVO_PAYLOAD 1
#endif
static const bool kPayload = SVO_PAYLOAD != 0;

namespace {

string filename;
uint64_t gridsize = 1024;           // main.cpp:30
uint64_t voxel_memory_limit = 2048;  // main.cpp:31
float sparseness_limit = 0.10f;      // main.cpp:32 (speed-only knob in the reference)
int color = SVO_COLOR_MODEL;
bool generate_levels = false;
bool verbose = false;
int device = 0;
bool write_trip = false;  // keep the reference's .trip/.tripdata partition files

struct TriInfo {
  string base;
  int version = 1;
  int geometry_only = 0;
  uint64_t n_triangles = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
};

void printInfo() {
  cout << "--------------------------------------------------------------------" << endl;
  if (kPayload)
    cout << "Out-Of-Core SVO Builder 1.6.4 (MI355X)" << endl;
  else
    cout << "Out-Of-Core SVO Builder 1.6.4 (MI355X) - Geometry only version" << endl;
  cout << "Linux " << endl;
  cout << "--------------------------------------------------------------------" << endl << endl;
}

void printHelp() {
  cout << "Example: svo_builder -f /home/jeroen/bunny.tri" << endl;
  cout << "" << endl;
  cout << "All available program options:" << endl;
  cout << "" << endl;
  cout << "-f <filename.tri>     Path to a .tri input file." << endl;
  cout << "-s <gridsize>         Voxel gridsize, should be a power of 2. Default 512." << endl;
  cout << "-l <memory_limit>     Memory limit for process, in Mb. Default 1024." << endl;
  cout << "-levels               Generate intermediary voxel levels by averaging voxel data" << endl;
  cout << "-c <option>           Coloring of voxels (Options: model (default), fixed, linear, normal)" << endl;
  cout << "-d <percentage>       Percentage of memory limit to be used additionaly for sparseness optimization" << endl;
  cout << "-v                    Be very verbose." << endl;
  cout << "-h                    Print help and exit." << endl;
}

void printInvalid() {
  cout << "Not enough or invalid arguments, please try again." << endl;
  cout << "At the bare minimum, I need a path to a .TRI file" << endl << "" << endl;
  printHelp();
}

bool isPowerOf2(unsigned int x) { return (x != 0) && !(x & (x - 1)); }

// main.cpp:100-196; returns false when the program must exit (status 0)
bool parseProgramParameters(int argc, char* argv[]) {
  string color_s = "Color from model (fallback to fixed color if model has no color)";
  cout << "Reading program parameters ..." << endl;
  if (argc < 3) {
    printInvalid();
    return false;
  }
  for (int i = 1; i < argc; i++) {
    const string a = argv[i];
    const bool has_next = i + 1 < argc;
    if (a == "-f" && has_next) {
      filename = argv[i + 1];
      if (filename.find(".tri") == string::npos) {
        cout << "Data filename does not end in .tri - I only support that file format" << endl;
        printInvalid();
        return false;
      }
      i++;
    } else if (a == "-s" && has_next) {
      gridsize = (uint64_t)atoi(argv[i + 1]);
      if (!isPowerOf2((unsigned int)gridsize)) {
        cout << "Requested gridsize is not a power of 2" << endl;
        printInvalid();
        return false;
      }
      i++;
    } else if (a == "-l" && has_next) {
      voxel_memory_limit = (uint64_t)atoi(argv[i + 1]);
      if (voxel_memory_limit <= 1) {
        cout << "Requested memory limit is nonsensical. Use a value >= 1" << endl;
        printInvalid();
        return false;
      }
      i++;
    } else if (a == "-d" && has_next) {
      int percent = atoi(argv[i + 1]);
      sparseness_limit = percent / 100.0f;
      if (sparseness_limit < 0) {
        cout << "Requested data memory limit is nonsensical. Use a value > 0" << endl;
        printInvalid();
        return false;
      }
      i++;
    } else if (a == "-v") {
      verbose = true;
    } else if (a == "-levels") {
      generate_levels = true;
    } else if (a == "-c" && has_next) {
      const string c = argv[i + 1];
      if (!kPayload) {
        cout << "You asked to generate colors, but we're only doing binary voxelisation." << endl;
      } else if (c == "model") {
        color = SVO_COLOR_MODEL;
      } else if (c == "linear") {
        color = SVO_COLOR_LINEAR;
        color_s = "Linear";
      } else if (c == "normal") {
        color = SVO_COLOR_NORMAL;
        color_s = "Normal";
      } else if (c == "fixed") {
        color = SVO_COLOR_FIXED;
        color_s = "Fixed";
      } else {
        cout << "Unrecognized color switch: " << c << ", so reverting to colors from model." << endl;
      }
      i++;
    } else if (a == "--device" && has_next) {
      device = atoi(argv[i + 1]);
      i++;
    } else if (a == "--write-trip") {
      write_trip = true;
    } else if (a == "-h") {
      printHelp();
      return false;
    } else {
      printInvalid();
      return false;
    }
  }
  if (verbose) {
    cout << "  filename: " << filename << endl;
    cout << "  gridsize: " << gridsize << endl;
    cout << "  memory limit: " << voxel_memory_limit << endl;
    cout << "  sparseness optimization limit: " << sparseness_limit << " resulting in "
         << (sparseness_limit * voxel_memory_limit) << " memory limit." << endl;
    cout << "  color type: " << color_s << endl;
    cout << "  generate levels: " << generate_levels << endl;
    cout << "  verbosity: " << verbose << endl;
  }
  return true;
}

bool file_exists(const string& n) {
  FILE* f = fopen(n.c_str(), "r");
  if (!f) return false;
  fclose(f);
  return true;
}

// tri_tools.h:75-115
int parseTriHeader(const string& fn, TriInfo& t) {
  std::ifstream file(fn.c_str());
  if (!file) {
    cout << "  Error: file " << fn << " does not exist." << endl;
    return 0;
  }
  t.base = fn.substr(0, fn.find_last_of("."));
  string line;
  file >> line;
  if (line != "#tri") {
    cout << "  Error: first line reads [" << line << "] instead of [#tri]" << endl;
    return 0;
  }
  file >> t.version;
  bool done = false;
  t.geometry_only = 0;
  while (file.good() && !done) {
    file >> line;
    if (line == "END") {
      done = true;
    } else if (line == "ntriangles") {
      file >> t.n_triangles;
    } else if (line == "geo_only") {
      file >> t.geometry_only;
    } else if (line == "bbox") {
      file >> t.bbox[0] >> t.bbox[1] >> t.bbox[2] >> t.bbox[3] >> t.bbox[4] >> t.bbox[5];
    } else {
      cout << "  unrecognized keyword [" << line << "], skipping" << endl;
      char c;
      do {
        c = (char)file.get();
      } while (file.good() && (c != '\n'));
    }
  }
  if (!done) {
    cout << "  error reading header" << endl;
    return 0;
  }
  return 1;
}

// The bbox as main.cpp sees it after the .trip header round trip (trip_tools.h:112,
// main.cpp:306-311): printed with 6 significant digits and parsed back.
float text_roundtrip(float v) {
  std::ostringstream o;
  o << v;
  std::istringstream i(o.str());
  float r = 0;
  i >> r;
  return r;
}

string val_to_string(uint64_t v) {
  std::ostringstream o;
  o << v;
  return o.str();
}

// The two output files the streamed build writes into (.octreenodes, .octreedata).
struct OutFiles {
  FILE* f[2] = {nullptr, nullptr};
};
int file_sink(void* user, int kind, const void* bytes, uint64_t size) {
  FILE* f = ((OutFiles*)user)->f[kind == SVO_SINK_DATA ? 1 : 0];
  return fwrite(bytes, 1, (size_t)size, f) == (size_t)size ? 0 : 1;
}

int die(svo_ctx* ctx, const char* what) {
  std::cerr << "svo_builder: " << what << " failed: " << svo_last_error() << endl;
  if (ctx) svo_ctx_destroy(ctx);
  return 1;
}

}  // namespace

int main(int argc, char* argv[]) {
  const auto t0 = std::chrono::steady_clock::now();
  printInfo();
  if (!parseProgramParameters(argc, argv)) return 0;

  TriInfo ti;
  cout << "Parsing tri header " << filename << " ..." << endl;
  if (parseTriHeader(filename, ti) != 1) return 0;
  if (!(file_exists(ti.base + ".tri") && file_exists(ti.base + ".tridata"))) {
    cout << "Not all required .tri or .tridata files exist. Please regenerate using tri_convert." << endl;
    return 0;
  }
  if (verbose) {
    cout << "  base_filename: " << ti.base << endl;
    cout << "  tri version: " << ti.version << endl;
    cout << "  geometry only: " << ti.geometry_only << endl;
    cout << "  n_triangles: " << ti.n_triangles << endl;
    cout << "  bbox min: " << ti.bbox[0] << " " << ti.bbox[1] << " " << ti.bbox[2] << endl;
    cout << "  bbox max: " << ti.bbox[3] << " " << ti.bbox[4] << " " << ti.bbox[5] << endl;
  }
  if (!kPayload && !ti.geometry_only) {
    cout << "You're using a .tri file which contains more than just geometry with a geometry-only SVO Builder! Regenerate that .tri file using tri_convert_binary." << endl;
    return 0;
  }
  if (kPayload && ti.geometry_only) {
    cout << "You're using a .tri file which contains only geometry with the regular SVO Builder! Regenerate that .tri file using tri_convert." << endl;
    return 0;
  }

  // stream the .tridata into HBM (the reference streams it through TriReader, 8192
  // triangles at a time)
  svo_ctx* ctx = nullptr;
  if (svo_ctx_create(device, &ctx) != SVO_OK) return die(nullptr, "svo_ctx_create");
  {
    uint64_t got = 0;
    const int rc = svo_mesh_load_file(ctx, (ti.base + ".tridata").c_str(), ti.n_triangles,
                                      kPayload ? 1 : 0, &got);
    if (rc == SVO_ERR_INVALID && got < ti.n_triangles) {
      std::cerr << "svo_builder: " << ti.base << ".tridata holds " << got
                << " triangles, the header says " << ti.n_triangles << endl;
      svo_ctx_destroy(ctx);
      return 1;
    }
    if (rc != SVO_OK) return die(ctx, "svo_mesh_load_file");
  }

  // PARTITIONING (main.cpp:293-302)
  cout << "Estimating best partition count ..." << endl;
  const uint64_t required = (gridsize * gridsize * gridsize) / 1024 / 1024;
  cout << "  to do this in-core I would need " << required << " Mb of system memory" << endl;
  const uint64_t n_partitions = svo_estimate_partitions(gridsize, voxel_memory_limit);
  if (n_partitions == 1)
    cout << "  memory limit of " << voxel_memory_limit << " Mb allows that" << endl;
  else
    cout << "  going to do it in " << n_partitions << " partitions of "
         << required / n_partitions << " Mb each." << endl;
  cout << "Partitioning data into " << n_partitions << " partitions ... ";
  cout.flush();
  const float u_part = (ti.bbox[3] - ti.bbox[0]) / (float)gridsize;  // partitioner.cpp:45
  std::vector<uint64_t> counts(n_partitions, 0);
  if (svo_partition(ctx, u_part, gridsize, n_partitions, counts.data()) != SVO_OK)
    return die(ctx, "svo_partition");
  cout << "done." << endl;
  const string base = ti.base + val_to_string(gridsize) + "_" + val_to_string(n_partitions);
  if (verbose)
    for (uint64_t i = 0; i < n_partitions; i++)
      cout << "  partition " << i << " - tri_count: " << counts[i] << endl;
  if (write_trip && svo_write_trip(ctx, base.c_str(), gridsize, ti.bbox, ti.version) != SVO_OK)
    return die(ctx, "svo_write_trip");

  // VOXELIZATION + SVO (main.cpp:311-387): the octree is streamed into the output files
  // partition by partition (OctreeBuilder.cpp:4-168 owns the two files the same way),
  // the copy of one chunk of records overlapping the write of the previous one
  const float unitlength =
      (text_roundtrip(ti.bbox[3]) - text_roundtrip(ti.bbox[0])) / (float)gridsize;
  const uint64_t morton_part = (gridsize * gridsize * gridsize) / n_partitions;
  OutFiles out;
  out.f[0] = fopen((base + ".octreenodes").c_str(), "wb");
  out.f[1] = fopen((base + ".octreedata").c_str(), "wb");
  if (!out.f[0] || !out.f[1]) {
    std::cerr << "svo_builder: cannot open output files for " << base << endl;
    svo_ctx_destroy(ctx);
    return 1;
  }
  if (svo_stream_begin(ctx, gridsize, n_partitions, generate_levels ? 1 : 0, file_sink, &out) !=
      SVO_OK)
    return die(ctx, "svo_stream_begin");
  uint64_t nfilled = 0;
  for (uint64_t i = 0; i < n_partitions; i++) {
    if (counts[i] == 0) continue;
    cout << "Voxelizing partition " << i << " ..." << endl;
    if (verbose)
      cout << "  reading " << counts[i] << " triangles from " << base << "_" << i << ".tripdata"
           << endl;
    uint64_t nf = 0;
    if (svo_voxelize(ctx, i, i * morton_part, (i + 1) * morton_part, unitlength, gridsize, color,
                     &nf) != SVO_OK)
      return die(ctx, "svo_voxelize");
    nfilled += nf;
    if (verbose) cout << "  found " << nf << " new voxels." << endl;
    cout << "Building SVO for partition " << i << " ..." << endl;
    if (svo_stream_partition(ctx, i) != SVO_OK) return die(ctx, "svo_stream_partition");
  }
  uint64_t n_nodes = 0, n_data = 0;
  int hazard = 0;
  if (svo_stream_end(ctx, &n_nodes, &n_data, &hazard) != SVO_OK) return die(ctx, "svo_stream_end");
  if (fclose(out.f[0]) != 0 || fclose(out.f[1]) != 0) {
    std::cerr << "svo_builder: writing the output files for " << base << " failed" << endl;
    svo_ctx_destroy(ctx);
    return 1;
  }
  if (hazard)
    std::cerr << "warning: the last filled voxel is the second-to-last Morton code (or the grid "
                 "is an empty 1^3): the reference's output is undefined here; the completed tree "
                 "is written"
              << endl;
  {  // octree_io.h:74-83
    std::ofstream h((base + ".octree").c_str());
    h << "#octreeheader 1" << endl;
    h << "gridlength " << gridsize << endl;
    h << "n_nodes " << n_nodes << endl;
    h << "n_data " << n_data << endl;
    h << "END" << endl;
  }
  cout << "done" << endl;
  cout << "Total amount of voxels: " << nfilled << endl;
  svo_ctx_destroy(ctx);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  cout << "Total MAIN time      : " << ms << " ms." << endl;
  return 0;
}
