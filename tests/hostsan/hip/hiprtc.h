#pragma once
#include <cstddef>
enum hiprtcResult { HIPRTC_SUCCESS = 0, HIPRTC_ERROR_COMPILATION = 6 };
typedef struct MockRtc* hiprtcProgram;
hiprtcResult hiprtcCreateProgram(hiprtcProgram* p, const char* src, const char* name, int n, const char** h,
                                 const char** hn);
hiprtcResult hiprtcCompileProgram(hiprtcProgram p, int n, const char** opts);
hiprtcResult hiprtcGetProgramLogSize(hiprtcProgram p, size_t* n);
hiprtcResult hiprtcGetProgramLog(hiprtcProgram p, char* log);
hiprtcResult hiprtcDestroyProgram(hiprtcProgram* p);
hiprtcResult hiprtcGetCodeSize(hiprtcProgram p, size_t* n);
hiprtcResult hiprtcGetCode(hiprtcProgram p, char* code);
const char* hiprtcGetErrorString(hiprtcResult r);
