// Test-infrastructure driver around the reference's damascene filter
// construction (damascene/src/textons/filters.cpp, compiled from the
// reference tree by `make ref`): the sPb derivative-of-Gaussian filter the way
// damascene/src/spectralPb/spectralPb.cu:706-712 builds it,
// gaussian_2D(sigma 1, 1, angle, deriv 1, no Hilbert, support 3, 3), 7 x 7
// floats, x-major (filters.cpp:264-341).  Used only by tests.
void gaussian_2D(float* filter, double sigma_x, double sigma_y, double ori, unsigned int deriv, bool hlbrt,
                 unsigned long support_x, unsigned long support_y);

extern "C" void ref_spb_filter(double ori, float* out) { gaussian_2D(out, 1, 1, ori, 1, false, 3, 3); }
