/* Twiddle constants exp(2 pi i m / L) for the in-register DFT sizes.
 * Generated once from 40-digit cos/sin values; each literal is the
 * correctly rounded double. */
#pragma once

template<int L> struct SmallTw;

template<> struct SmallTw<3> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[3] = {1.0, -0.5, -0.5}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[3] = {0.0, 0.8660254037844386, -0.8660254037844386}; return t[m]; }
};
template<> struct SmallTw<5> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475, 0.30901699437494745}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731, -0.9510565162951535}; return t[m]; }
};
template<> struct SmallTw<6> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[6] = {0.0, 0.8660254037844386, 0.8660254037844386, 4.1340642196527976e-43, -0.8660254037844386, -0.8660254037844386}; return t[m]; }
};
template<> struct SmallTw<8> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[8] = {1.0, 0.7071067811865476, 2.0670321098263988e-43, -0.7071067811865476, -1.0, -0.7071067811865476, 2.2338764406549882e-41, 0.7071067811865476}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[8] = {0.0, 0.7071067811865476, 1.0, 0.7071067811865476, 4.1340642196527976e-43, -0.7071067811865476, -1.0, -0.7071067811865476}; return t[m]; }
};
template<> struct SmallTw<10> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[10] = {1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475, -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[10] = {0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731, 4.1340642196527976e-43, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731}; return t[m]; }
};
template<> struct SmallTw<12> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[12] = {1.0, 0.8660254037844386, 0.5, 2.0670321098263988e-43, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 2.2338764406549882e-41, 0.5, 0.8660254037844386}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[12] = {0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5, 4.1340642196527976e-43, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5}; return t[m]; }
};
template<> struct SmallTw<16> {
  static __device__ __forceinline__ double re(int m) { constexpr double t[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 2.0670321098263988e-43, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 2.2338764406549882e-41, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867}; return t[m]; }
  static __device__ __forceinline__ double im(int m) { constexpr double t[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 4.1340642196527976e-43, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898}; return t[m]; }
};
