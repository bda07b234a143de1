/*
 * gpb_ucm.hpp — C++ entry points of libgpb.so under the names, argument
 * order and output layouts of the yanyp/UCM-MPI-GPU gPb calls, so that
 * UCMgenerator-style callers switch by relinking.
 *
 * The reference passes OpenCV matrices (cv::Mat, CV_64F).  OpenCV is not a
 * dependency of this library, so the CPU-side entry points take gpb_ucm::Mat,
 * a row-major double matrix with the members those functions use (rows,
 * cols, channels, data, at, ptr); converting a cv::Mat is a copy of
 * rows * cols * channels doubles.  computeGPb / getCudaDeviceCount keep the
 * reference's exact signatures (global namespace, C++ linkage).
 *
 * Every function runs on the GPU through the C ABI (gpb_capi.h); there is no
 * CPU fallback.  Failures throw std::runtime_error (std::invalid_argument for
 * bad arguments) from the gpb_ucm functions, and int64_t(1) from computeGPb,
 * which is what the reference's CUDA error check throws
 * (damascene/include/exception.h:16-28) and what its caller catches
 * (src/globalPb.cpp:283-286).
 */
#ifndef GPB_UCM_HPP
#define GPB_UCM_HPP

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace gpb_ucm {

/* Row-major matrix of doubles with interleaved channels (the CV_64FC1 /
 * CV_64FC3 matrices of the reference). */
struct Mat {
  int rows = 0, cols = 0, channels = 1;
  std::vector<double> data;
  Mat() = default;
  Mat(int r, int c, int ch = 1, double v = 0.0) : rows(r), cols(c), channels(ch), data(size_t(r) * c * ch, v) {}
  double& at(int r, int c, int ch = 0) { return data[(size_t(r) * cols + c) * channels + ch]; }
  double at(int r, int c, int ch = 0) const { return data[(size_t(r) * cols + c) * channels + ch]; }
  double* ptr(int r) { return data.data() + size_t(r) * cols * channels; }
  const double* ptr(int r) const { return data.data() + size_t(r) * cols * channels; }
  bool empty() const { return data.empty(); }
};

struct Size {
  int width = 0, height = 0;
};

/* The 12 local cues x 8 orientations and the texton map of det_mPb
 * (include/detmpb.h:8-45).  Owns its matrices. */
class detmpb {
 public:
  detmpb() = default;
  ~detmpb();
  detmpb(const detmpb&) = delete;
  detmpb& operator=(const detmpb&) = delete;
  std::vector<Mat*>* GetBg1() { return &cue_[0]; }
  std::vector<Mat*>* GetBg2() { return &cue_[1]; }
  std::vector<Mat*>* GetBg3() { return &cue_[2]; }
  std::vector<Mat*>* GetCga1() { return &cue_[3]; }
  std::vector<Mat*>* GetCga2() { return &cue_[4]; }
  std::vector<Mat*>* GetCga3() { return &cue_[5]; }
  std::vector<Mat*>* GetCgb1() { return &cue_[6]; }
  std::vector<Mat*>* GetCgb2() { return &cue_[7]; }
  std::vector<Mat*>* GetCgb3() { return &cue_[8]; }
  std::vector<Mat*>* GetTg1() { return &cue_[9]; }
  std::vector<Mat*>* GetTg2() { return &cue_[10]; }
  std::vector<Mat*>* GetTg3() { return &cue_[11]; }
  Mat* GetTextons() { return textons_; }
  /* Takes ownership of 12 x 8 cue maps (bg1..3, cga1..3, cgb1..3, tg1..3)
   * and the texton map, releasing what it held (the role of
   * detmpb::AssignPtrsUsingShallowCopy, src/detmpb.cpp:55-127). */
  void Assign(std::vector<Mat*> (&cues)[12], Mat* textons);
  void clear();

 private:
  std::vector<Mat*> cue_[12];
  Mat* textons_ = nullptr;
};

/* Local cues of a 3-channel image with values in [0, 1] (channel 0 = red,
 * 1 = green, 2 = blue; each value is rounded to the nearest multiple of
 * 1/255): fills the 12 cue vectors of 8 orientation maps (H x W) and the
 * texton map.  Replaces det_mPb (src/det_mPb.cpp:9-22). */
void det_mPb(detmpb& returnVals, Mat& im);

/* mPb = weighted sum of the 12 cues, max over orientations, oriented nonmax
 * (tolerance pi/8) and the clip of src/multiscalePb.cpp:251-261.  Allocates
 * mPb_nmax (new Mat, H x W); mPb_nmax_rsz aliases it (rsz >= 1, as in the
 * reference; rsz < 1 is rejected).  nChan must be 3.  outFile is unused, as
 * in the reference.  Replaces multiscalePb (src/multiscalePb.cpp:16-283). */
void multiscalePb(Mat*& mPb_nmax, Mat*& mPb_nmax_rsz, detmpb& detmpb_obj, Mat& imd, Size orig_size, int nChan,
                  std::string outFile, double rsz = 1.0);

/* sPb: intervening-contour W of mPb, the nvec smallest generalised
 * eigenpairs of (D - W, D) to residual 1e-3, 8 oriented sPb maps (new Mat
 * each, H x W) from eigenvectors 1..nvec-1.  Replaces spectralPb
 * (src/spectralPb.cpp:33-343); nvec <= 32. */
void spectralPb(std::vector<Mat*>& sPb, Mat* mPb_rsz, Size orig_size, std::string outFile2, int nvec = 17);

/* The affinity of the lattice l1 ((H+1) x W, mPb in rows 1..H) and l2
 * (H x (W+1), mPb in columns 1..W) as buildW exports it: valD[N] = row sums
 * of W (W_pp = 1), and (I, J, val) = the entries with J >= I (the diagonal
 * and upper triangle), pixels numbered in scanline order (y * W + x), rows
 * ascending, columns ascending within a row.  l1 and l2 must hold the same
 * mPb map, as spectralPb builds them.  Replaces buildW
 * (src/buildW.cpp:23-246). */
void buildW(std::vector<double>& valD, std::vector<double>& val, std::vector<int>& I, std::vector<int>& J, Mat& l1,
            Mat& l2);

/* Whole gPb of the binary PPM (P6, 8-bit) image imgFile: gPb[o] = the 8
 * oriented gPb maps before calibration, o = split angle o * pi / 8.  Uses
 * 17 eigenvectors, as the reference's GlobalPb does (src/globalPb.cpp:65).
 * When outFile is not empty, writes outFile + "_gPb.yml" (gPb_0 .. gPb_7)
 * and outFile + "_mPb.yml" (mPb_rsz_mat) in OpenCV's YAML matrix format
 * (src/globalPb.cpp:50-53, 113-121); the mPb file is output only (the
 * reference's read-back path, :42-49, leaves its cue set empty).  Replaces
 * GlobalPb (src/globalPb.cpp:15-126); rsz must be 1. */
void GlobalPb(std::string imgFile, std::string outFile, std::vector<Mat>& gPb, double rsz = 1.0);

/* Device for the gpb_ucm functions (default 0). */
void set_device(int device);

/* Texton codebook (13 x 32 floats, filter-major) used by det_mPb, GlobalPb
 * and computeGPb; nullptr (default) = one per image by seeded k-means++
 * (seed 1, 10 Lloyd iterations) on the image's own Schmid responses. */
void set_texton_codebook(const float* codebook);

/* Learned warm start for computeGPb (no counterpart in the reference):
 * params = gpb_warm_param_count() floats (e.g. the trained
 * ucm-mpi-gpu_amd/weights/warm_start.f32), copied; later computeGPb calls
 * start block Lanczos from the network's block, with plain Lanczos as the
 * fallback.  nullptr (default) = plain Lanczos. */
void set_warm_start(const float* params);

/* Binary PPM (P6, maxval 255) reader: rgb = H x W x 3 bytes. */
bool read_ppm(const std::string& path, int& width, int& height, std::vector<uint8_t>& rgb);

/* Writes single-channel matrices in OpenCV's FileStorage YAML format
 * (%YAML:1.0, "name: !!opencv-matrix" with rows, cols, dt: d, data). */
void write_opencv_yaml(const std::string& path, const std::vector<std::pair<std::string, const Mat*>>& mats);

/* The byte frame of a vector of CV_64F matrices (gpb_frame_mats): what
 * MpiSendMatVec sends (src/utilities.cpp:415-438), in one buffer. */
std::vector<uint8_t> frame_mat_vec(const std::vector<Mat>& mats);
/* Inverse of frame_mat_vec (MpiRecvMatVec, src/utilities.cpp:440-460);
 * throws std::invalid_argument on a malformed frame or a non-CV_64F matrix. */
std::vector<Mat> unframe_mat_vec(const std::vector<uint8_t>& frame);

/* The ultrametric contour map of the 8 oriented gPb maps (each H x W, the
 * order GlobalPb returns them in): (2H+1) x (2W+1) when fmt is "doubleSize",
 * else (H+1) x (W+1).  pieceFilePat names the piece in the reference (its
 * debug images); it is unused here.  Replaces Contours2Ucm
 * (src/contours2ucm.cpp:405-490; gpb_contours2ucm). */
Mat Contours2Ucm(const std::string& pieceFilePat, const std::vector<Mat>& pb_oriented, const std::string& fmt);

}  // namespace gpb_ucm

/* Whole gPb of one packed image on GPU rank % getCudaDeviceCount():
 * data = H*W words 0x00BBGGRR (r in bits 0-7); *hostGPb = malloc'd H*W
 * floats (max over orientations, 0 where mPb_nmax <= 0.05); *hostGPbAllConcat
 * = malloc'd 8 blocks of W*H floats, block i column-major (index x*H + y),
 * the calibration sigmoid of the oriented gPb at split angle (7 - i) * pi / 8:
 * the reference's oriMap {4,5,6,7,0,1,2,3} (damascene.cu:598-603) over filter
 * slots s that hold the split angle ((3 - s) mod 8) * pi / 8 (the filter for
 * pi/2 + (7 - s) pi/8 applied row-major, spectralPb.cu:101,706-712).  The
 * caller frees both with free().  9 eigenvectors to
 * residual 1e-3 (damascene/src/damascene.cu:307-309).  Calls on one device
 * are serialised.  Replaces computeGPb (damascene/include/globalPb.h:37,
 * damascene/src/damascene.cu:249-661). */
void computeGPb(unsigned int rank, unsigned int width, unsigned int height, unsigned int* data, float** hostGPb,
                float** hostGPbAllConcat);

/* Number of visible GPUs.  Replaces getCudaDeviceCount
 * (damascene/include/globalPb.h:44, damascene/src/damascene.cu:242-247). */
int getCudaDeviceCount();

#endif /* GPB_UCM_HPP */
