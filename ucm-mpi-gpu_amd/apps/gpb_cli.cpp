// gpb_cli: command-line front end of libgpb.so through the reference-named
// C++ entry points (gpb_ucm.hpp).
//
//   gpb_cli gpb <in.ppm> <outprefix>    GlobalPb: <outprefix>_gPb.yml and _mPb.yml
//   gpb_cli compute <in.ppm> <out.bin>  computeGPb: hostGPb then hostGPbAllConcat (float32)
//   gpb_cli shims <in.ppm> <outdir>     every entry point, raw dumps for cross-checks
//   gpb_cli yaml <out.yml>              a 2 x 3 matrix in the YAML format (format check)
//   gpb_cli devices                     getCudaDeviceCount()
//
// Exit status 0 on success, 2 on bad usage or input, 3 on a library error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpb_capi.h"
#include "gpb_ucm.hpp"

using gpb_ucm::Mat;

namespace {

template <typename T>
void dump(const std::string& path, const T* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  if (n && std::fwrite(p, sizeof(T), n, f) != n) {
    std::fclose(f);
    throw std::runtime_error("short write " + path);
  }
  std::fclose(f);
}

void dump_mats(const std::string& path, const std::vector<Mat*>& v) {
  std::vector<double> all;
  for (const Mat* m : v) all.insert(all.end(), m->data.begin(), m->data.end());
  dump(path, all.data(), all.size());
}

Mat read_image(const std::string& path) {
  int W, H;
  std::vector<uint8_t> rgb;
  if (!gpb_ucm::read_ppm(path, W, H, rgb)) throw std::invalid_argument("cannot read binary PPM " + path);
  Mat im(H, W, 3);
  for (size_t i = 0; i < rgb.size(); i++) im.data[i] = rgb[i] / 255.0;
  return im;
}

std::vector<unsigned int> packed_image(const std::string& path, int& W, int& H) {
  std::vector<uint8_t> rgb;
  if (!gpb_ucm::read_ppm(path, W, H, rgb)) throw std::invalid_argument("cannot read binary PPM " + path);
  std::vector<uint32_t> p(size_t(W) * H);
  gpb_pack_rgb(rgb.data(), p.size(), p.data());
  return std::vector<unsigned int>(p.begin(), p.end());
}

int run_shims(const std::string& in, const std::string& dir) {
  Mat im = read_image(in);
  const int H = im.rows, W = im.cols;
  {  // det_mPb
    gpb_ucm::detmpb d;
    gpb_ucm::det_mPb(d, im);
    std::vector<Mat*> all;
    for (auto* v : {d.GetBg1(), d.GetBg2(), d.GetBg3(), d.GetCga1(), d.GetCga2(), d.GetCga3(), d.GetCgb1(),
                    d.GetCgb2(), d.GetCgb3(), d.GetTg1(), d.GetTg2(), d.GetTg3()})
      all.insert(all.end(), v->begin(), v->end());
    dump_mats(dir + "/cues.f64", all);
    dump_mats(dir + "/textons.f64", {d.GetTextons()});
  }
  Mat* nmax = nullptr;
  {  // multiscalePb
    gpb_ucm::detmpb d;
    Mat* rsz = nullptr;
    gpb_ucm::multiscalePb(nmax, rsz, d, im, gpb_ucm::Size{W, H}, 3, "");
    if (rsz != nmax) throw std::runtime_error("mPb_nmax_rsz must alias mPb_nmax at rsz = 1");
    dump_mats(dir + "/nmax.f64", {nmax});
  }
  for (int nvec : {9, 17}) {  // spectralPb
    std::vector<Mat*> spb;
    gpb_ucm::spectralPb(spb, nmax, gpb_ucm::Size{W, H}, "", nvec);
    dump_mats(dir + "/spb" + std::to_string(nvec) + ".f64", spb);
    for (Mat* m : spb) delete m;
  }
  {  // buildW on the lattice spectralPb builds (spectralPb.cpp:55-59)
    Mat l1(H + 1, W), l2(H, W + 1);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) l1.at(y + 1, x) = l2.at(y, x + 1) = nmax->at(y, x);
    std::vector<double> valD, val;
    std::vector<int> I, J;
    gpb_ucm::buildW(valD, val, I, J, l1, l2);
    dump(dir + "/valD.f64", valD.data(), valD.size());
    dump(dir + "/val.f64", val.data(), val.size());
    dump(dir + "/I.i32", I.data(), I.size());
    dump(dir + "/J.i32", J.data(), J.size());
  }
  delete nmax;
  {  // GlobalPb
    std::vector<Mat> g;
    gpb_ucm::GlobalPb(in, dir + "/img", g);
    std::vector<Mat*> ptrs;
    for (auto& m : g) ptrs.push_back(&m);
    dump_mats(dir + "/globalpb.f64", ptrs);
  }
  {  // computeGPb
    int w, h;
    std::vector<unsigned int> data = packed_image(in, w, h);
    float *g = nullptr, *all = nullptr;
    computeGPb(0, unsigned(w), unsigned(h), data.data(), &g, &all);
    dump(dir + "/gmax.f32", g, size_t(w) * h);
    dump(dir + "/gall.f32", all, size_t(w) * h * 8);
    std::free(g);
    std::free(all);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  try {
    if (mode == "devices" && argc == 2) {
      std::printf("%d\n", getCudaDeviceCount());
      return 0;
    }
    if (mode == "yaml" && argc == 3) {
      Mat m(2, 3);
      for (int i = 0; i < 6; i++) m.data[size_t(i)] = 0.5 * i - 1.0 / 3.0;
      gpb_ucm::write_opencv_yaml(argv[2], {{"m", &m}});
      return 0;
    }
    if (mode == "gpb" && argc == 4) {
      std::vector<Mat> g;
      gpb_ucm::GlobalPb(argv[2], argv[3], g);
      return 0;
    }
    if (mode == "compute" && argc == 4) {
      int w, h;
      std::vector<unsigned int> data = packed_image(argv[2], w, h);
      float *g = nullptr, *all = nullptr;
      computeGPb(0, unsigned(w), unsigned(h), data.data(), &g, &all);
      FILE* f = std::fopen(argv[3], "wb");
      if (!f) throw std::runtime_error("cannot write output");
      std::fwrite(g, sizeof(float), size_t(w) * h, f);
      std::fwrite(all, sizeof(float), size_t(w) * h * 8, f);
      std::fclose(f);
      std::free(g);
      std::free(all);
      return 0;
    }
    if (mode == "shims" && argc == 4) return run_shims(argv[2], argv[3]);
    if (mode == "ucm" && (argc == 6 || argc == 7)) {
      // 8 oriented H x W maps (raw doubles) -> Contours2Ucm -> raw doubles
      const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
      if (w < 2 || h < 2) throw std::invalid_argument("bad map size");
      std::vector<Mat> pbo(8, Mat(h, w));
      FILE* f = std::fopen(argv[2], "rb");
      if (!f) throw std::invalid_argument("cannot read maps");
      for (auto& m : pbo)
        if (std::fread(m.data.data(), sizeof(double), m.data.size(), f) != m.data.size()) {
          std::fclose(f);
          throw std::invalid_argument("short maps file");
        }
      std::fclose(f);
      const Mat u = gpb_ucm::Contours2Ucm("piece", pbo, argc == 7 ? argv[6] : "");
      FILE* o = std::fopen(argv[5], "wb");
      if (!o) throw std::runtime_error("cannot write output");
      std::fwrite(u.data.data(), sizeof(double), u.data.size(), o);
      std::fclose(o);
      return 0;
    }
  } catch (const std::invalid_argument& e) {
    std::fprintf(stderr, "gpb_cli: %s\n", e.what());
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gpb_cli: %s\n", e.what());
    return 3;
  } catch (int64_t) {
    std::fprintf(stderr, "gpb_cli: computeGPb failed\n");
    return 3;
  }
  std::fprintf(stderr,
               "usage: gpb_cli gpb <in.ppm> <outprefix> | compute <in.ppm> <out.bin> | shims <in.ppm> <outdir> |"
               " ucm <maps.f64> <W> <H> <out.f64> [doubleSize] | yaml <out.yml> | devices\n");
  return 2;
}
