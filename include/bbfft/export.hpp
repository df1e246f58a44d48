// Symbol visibility for the MI355X build of the bbfft API.
#ifndef BBFFT_EXPORT_HPP
#define BBFFT_EXPORT_HPP

#define BBFFT_EXPORT __attribute__((visibility("default")))

#endif // BBFFT_EXPORT_HPP
