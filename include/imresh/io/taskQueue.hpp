// imresh::io task queue -- drop-in for the reference declarations
// (src/imresh/io/taskQueue.hpp:38-84).  addTask runs the reconstruction on
// an MI355X stream slot and then calls writeOutFunc(ioData, W, H, fileName)
// on a worker thread, like the reference.  Parameter defaults that are <= 0
// (or 0 for counts) follow cudaShrinkWrap (cudaShrinkWrap.cu:193-200), which
// the reference queue calls.  addTask before taskQueueInit (the reference
// asserts, taskQueue.cu:163) drops the task without calling writeOutFunc;
// the C entry point imresh_add_task reports IMRESH_ERR_NOT_INITIALIZED.
#pragma once

#include <functional>
#include <string>

namespace imresh
{
namespace io
{

    using WriteOutCallBack = std::function<
                                 void ( float *,
                                        unsigned int const,
                                        unsigned int const,
                                        std::string )
                             >;

    void addTask(
        float * const ioData,
        unsigned int const imageWidth                   ,
        unsigned int const imageHeight                  ,
        WriteOutCallBack writeOutFunc                   ,
        std::string      fileName                       ,
        unsigned int     nCycles                  = 20  ,
        unsigned int     nHioCycles               = 20  ,
        float            targetError              = 1e-5,
        float            hioBeta                  = 0.9f,
        float            intensityCutOffAutoCorel = 0.04,
        float            intensityCutOff          = 0.2 ,
        float            sigma0                   = 3.0 ,
        float            sigmaChange              = 0.01
    );

    void taskQueueInit( );

    void taskQueueDeinit( );

} // namespace io
} // namespace imresh
