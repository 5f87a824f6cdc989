/* Operator plugin ABI: the six symbols a runtime-compiled operator library
 * exports, as the reference's generated plugin source defines them
 * (PySparQ/pysparq/dynamic_operator/compiler.py:41-105) and its loaders bind them
 * (operator_wrapper.py:86-160 through ctypes; dynamic_operator_loader.cpp:54,131
 * through dlopen(RTLD_NOW | RTLD_LOCAL) + dlsym).
 *
 * The types are the host C++ classes of this engine (sparqsim_amd/csrc/host/
 * sparq.hpp, re-exported under namespace qram_simulator by basic_components.h).
 * A plugin is built with g++ against libsparq_host.so, so it shares the process-
 * global register table with the Python module.  Null arguments are ignored and
 * there is no error return, as in the reference; where the reference lets a C++
 * exception escape (terminating the process), the generated apply_operator /
 * apply_operator_dag catch it and park the message, which the Python proxy
 * raises as RuntimeError.
 *
 * In C these are declared over opaque pointers; a C++ plugin defines them with
 * the class types (BaseOperator*, SparseState*), which have the same ABI. */
#ifndef SPARQ_PLUGIN_H
#define SPARQ_PLUGIN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sparq_plugin_operator sparq_plugin_operator; /* qram_simulator::BaseOperator */
typedef struct sparq_plugin_state sparq_plugin_state;       /* qram_simulator::SparseState  */

/* compiler.py:49-51 -- constructs the plugin's class; the parameter list is the
 * user's constructor signature (C scalar types), here shown without arguments. */
typedef sparq_plugin_operator* (*sparq_create_operator_fn)();

/* compiler.py:53-55 */
void destroy_operator(sparq_plugin_operator* op);

/* compiler.py:57-59 -- the class name. */
const char* get_operator_name(void);

/* compiler.py:64-70 -- op(*state): the engine downloads the device state into a
 * std::vector<System>, runs the plugin's operator() and uploads the result. */
void apply_operator(sparq_plugin_operator* op, sparq_plugin_state* state);

/* compiler.py:72-78 -- op->dag(*state). */
void apply_operator_dag(sparq_plugin_operator* op, sparq_plugin_state* state);

/* compiler.py:80-83 -- "BaseOperator" or "SelfAdjointOperator". */
const char* get_base_class(void);

#ifdef __cplusplus
}
#endif
#endif /* SPARQ_PLUGIN_H */
