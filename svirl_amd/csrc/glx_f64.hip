// FP64 instantiation of every launcher. Built with -ffp-contract=off so the
// arithmetic keeps the reference's separate multiply and add roundings.
#include "glx_launch_impl.hpp"

template struct glx::L<double>;
