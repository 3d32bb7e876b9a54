/** \file buffer.hh
 * \brief comm_buffer of the drop-in API (reference: buffer.hh:16-44).
 *
 * The reference shares one grow-only host buffer between the levels of a
 * rank for MPI packing and the LU right-hand side.  On the GPU every level's
 * working storage is allocated with the hierarchy in HBM, so the buffer only
 * keeps its type and methods for source compatibility: the multigrid
 * constructor still takes one, and check_buf* still guarantee host space. */
#ifndef PGMG_AMD_BUFFER_HH
#define PGMG_AMD_BUFFER_HH

#include <cstdlib>

#include "common.hh"

class comm_buffer {
	public:
		/** Host scratch, grown on demand. */
		double *buf;
		/** Its size in bytes. */
		size_t mem;
		comm_buffer() : buf(NULL), mem(0) {}
		~comm_buffer() {if (mem > 0) free(buf);}
		inline void check_buf(int q) {reallocate(q * sizeof(double));}
		inline void check_buf_int(int q) {reallocate(q * sizeof(int));}
		inline void check_buf_uint(int q) {reallocate(q * sizeof(unsigned int));}
		inline void check_buf_float(int q) {reallocate(q * sizeof(float));}
	private:
		void reallocate(size_t nmem) {
			if (nmem <= mem) return;
			if (mem > 0) free(buf);
			mem = nmem;
			buf = (double *)malloc(mem);
			if (buf == NULL) p_fatal_error("Buffer allocation failed", PGMG_MEMORY_ERROR);
		}
		comm_buffer(const comm_buffer &);
		comm_buffer &operator=(const comm_buffer &);
};

#endif
