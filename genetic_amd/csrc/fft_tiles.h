// Tile shapes of the strided FFT passes (threads per block, adjacent kz
// columns per tile) per precision and size.  Written by
// genetic_amd/tune_fft.py from timings of every candidate shape on the GPU;
// sizes without an entry take the default.
#pragma once

namespace gic {

// The y passes (lines 4 KB apart) where they take a shape of their own:
// declared only for the sizes listed after the primary template.
template <typename T, int N> struct YTileShape;
// float n=1024: 16-column tiles (128-byte segments, one block per CU) took
// the step's y passes from 36.8 to 34.8 ms against the x passes' 8 columns
template <> struct YTileShape<float, 1024> {
  static constexpr int threads = 1024;
  static constexpr int cols = 16;
};

template <typename T, int N> struct TileShape {
  static constexpr int threads = 512;
  static constexpr int cols = sizeof(T) == 4 ? 8 : 4;
};
// float n=128: 0.1124 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<float, 128> {
  static constexpr int threads = 256;
  static constexpr int cols = 8;
};
// float n=256: 0.3011 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<float, 256> {
  static constexpr int threads = 512;
  static constexpr int cols = 16;
};
// float n=512: 2.9122 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<float, 512> {
  static constexpr int threads = 1024;
  static constexpr int cols = 16;
};
// float n=1024: 23.2576 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<float, 1024> {
  static constexpr int threads = 1024;
  static constexpr int cols = 8;
};
// double n=64: 0.0874 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<double, 64> {
  static constexpr int threads = 512;
  static constexpr int cols = 8;
};
// double n=128: 0.1455 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<double, 128> {
  static constexpr int threads = 256;
  static constexpr int cols = 4;
};
// double n=256: 0.6072 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<double, 256> {
  static constexpr int threads = 256;
  static constexpr int cols = 8;
};
// double n=512: 5.5496 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<double, 512> {
  static constexpr int threads = 256;
  static constexpr int cols = 8;
};
// double n=1024: 47.7723 ms for the four timed transforms (round trip, ZA, windowed c2r, windowed r2c)
template <> struct TileShape<double, 1024> {
  static constexpr int threads = 1024;
  static constexpr int cols = 8;
};

}  // namespace gic
