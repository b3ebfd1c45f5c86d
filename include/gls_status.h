/* Status codes returned by every libgls entry point.  The reference prints
 * and calls exit(-1) on every failure (src/PreProcess.cpp:44-49,
 * src/simulate.cu:1359-1365); the library instead returns one of these and
 * keeps a thread-local message for gls_last_error(). */
#ifndef GLS_STATUS_H
#define GLS_STATUS_H

#define GLS_OK 0
#define GLS_E_ARG (-1)         /* bad argument / handle */
#define GLS_E_FORMAT (-2)      /* malformed vlib / gv / SDF / VCD / inter.dat */
#define GLS_E_OOM (-3)         /* device or host memory exhausted */
#define GLS_E_DEVICE (-4)      /* HIP runtime error or no usable device */
#define GLS_E_UNSUPPORTED (-5) /* netlist outside the device kernel's domain */
#define GLS_E_CANCELLED (-6)   /* job cancelled or past its deadline */
#define GLS_E_RANGE (-7)       /* time or size outside the 30-bit event word */
#define GLS_E_STATE (-8)       /* call made in the wrong state */
#define GLS_E_IO (-9)          /* file cannot be opened / written */
#define GLS_E_CYCLE (-10)      /* combinational loop or undriven input net */

#endif
