# CMake package of the MI355X build of linkcell, usable in place:
#   find_package(linkcell CONFIG REQUIRED PATHS <repo>/linkcell_amd/cmake)
# It provides the targets of the reference package (linkcellConfig.cmake.in,
# CMakeLists.txt:144-163) over the one HIP library of this build, which
# exports both the C waist (lc_knearest*) and the device waist (lc_gpu_*):
#   linkcell::linkcell  the shared library and include/
#   linkcell::cpp       linkcell::linkcell + C++17 (include/linkcell.hpp)
#   linkcell::gpu       linkcell::linkcell + LINKCELL_HAS_GPULITE / LINKCELL_HAS_HIP
#                       (include/linkcell_gpu.h, include/linkcell_gpu.hpp)
get_filename_component(_linkcell_pkg "${CMAKE_CURRENT_LIST_DIR}/.." ABSOLUTE)
get_filename_component(_linkcell_root "${_linkcell_pkg}/.." ABSOLUTE)
set(LINKCELL_INCLUDE_DIR "${_linkcell_root}/include")
set(LINKCELL_LIBRARY "${_linkcell_pkg}/liblinkcell_amd.so")
if(NOT EXISTS "${LINKCELL_LIBRARY}")
  set(linkcell_FOUND FALSE)
  set(linkcell_NOT_FOUND_MESSAGE
      "${LINKCELL_LIBRARY} is not built (python -m linkcell_amd.build)")
  return()
endif()

if(NOT TARGET linkcell::linkcell)
  add_library(linkcell::linkcell SHARED IMPORTED)
  set_target_properties(linkcell::linkcell PROPERTIES
    IMPORTED_LOCATION "${LINKCELL_LIBRARY}"
    IMPORTED_NO_SONAME TRUE
    INTERFACE_INCLUDE_DIRECTORIES "${LINKCELL_INCLUDE_DIR}")
endif()

if(NOT TARGET linkcell::cpp)
  add_library(linkcell::cpp INTERFACE IMPORTED)
  set_target_properties(linkcell::cpp PROPERTIES
    INTERFACE_LINK_LIBRARIES linkcell::linkcell
    INTERFACE_COMPILE_FEATURES cxx_std_17)
endif()

if(NOT TARGET linkcell::gpu)
  add_library(linkcell::gpu INTERFACE IMPORTED)
  set_target_properties(linkcell::gpu PROPERTIES
    INTERFACE_LINK_LIBRARIES linkcell::linkcell
    INTERFACE_COMPILE_DEFINITIONS "LINKCELL_HAS_GPULITE=1;LINKCELL_HAS_HIP=1"
    INTERFACE_COMPILE_FEATURES cxx_std_17)
endif()

set(linkcell_FOUND TRUE)
