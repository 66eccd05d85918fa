// jms_sizes.h -- line lengths (grid edges nx, ny) with compiled kernels.
// Every length has an FFT plan in jms_fft_device.hpp (PlanFor<N>); the build
// compiles one translation unit per length (jms_kernels_size.hip, -DJMS_N=N).
#pragma once
#define JMS_SIZES(X)                                                                   \
  X(16) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(320) X(360) X(384) X(480) \
  X(512) X(625) X(640) X(720) X(768) X(1024) X(2048)
