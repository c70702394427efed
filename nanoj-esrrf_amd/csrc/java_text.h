// java_text.h -- the text the reference's Java code builds constants from.
//
// LiveSRRF_CL bakes its constants into the kernel source as text: `"" + value` renders a
// Java float with Float.toString, and the OpenCL compiler reads that text back as a
// double literal that the kernel casts to float.  For most constants the round trip
// returns the float itself, but `"" + 2 * sigma + 1` (LiveSRRF_CL.java:318) is string
// concatenation: Float.toString(2*sigma) followed by the character '1'.
#pragma once

#include <string>

namespace esrrf {

// java.lang.Float.toString(x): the shortest decimal that rounds to x, the closest to x
// among several (two-digit decimals compete when one digit suffices), written as
// "ddd.ddd" for 1e-3 <= |x| < 1e7 and as "d.dddE[-]n" otherwise, with at least one
// digit after the point.
std::string java_float_to_string(float x);

// (float) of the OpenCL double literal `text`: correctly rounded decimal -> double ->
// float, independent of the process locale.
float opencl_literal_to_float(const std::string &text);

// The tSO the BIC and 3-D kernels compare against: (float)(Float.toString(2*sigma) + "1").
float java_two_sigma_plus_one(float two_sigma);

}  // namespace esrrf
