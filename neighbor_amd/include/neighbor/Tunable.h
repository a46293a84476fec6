// Tunable.h -- the single tunable launch parameter of the reference's objects
// (include/neighbor/Tunable.h:28-133): a set of valid values, LaunchParameters
// {tunable, stream} and checkParameter(), which throws std::runtime_error for a
// value outside the set.
#ifndef NEIGHBOR_TUNABLE_H_
#define NEIGHBOR_TUNABLE_H_

#include "Runtime.h"

#include <set>
#include <stdexcept>
#include <vector>

namespace neighbor
{

//! Raised by checkParameter for a value outside the tunable set.
class tuning_error : public std::runtime_error
    {
    public:
        tuning_error() : std::runtime_error("Invalid tuning parameter.") {}
    };

template<typename T>
class Tunable
    {
    public:
        Tunable(T first, T last, T step)
            {
            for (T v = first; v <= last; v += step) m_valid.insert(v);
            }

        explicit Tunable(const std::vector<T>& values)
            {
            setTunableParameters(values);
            }

        std::vector<T> getTunableParameters() const
            {
            return std::vector<T>(m_valid.begin(), m_valid.end());
            }

        void setTunableParameters(const std::vector<T>& values)
            {
            m_valid = std::set<T>(values.begin(), values.end());
            }

        struct LaunchParameters
            {
            typedef T type;
            explicit LaunchParameters(T tunable_) : tunable(tunable_), stream(0) {}
            LaunchParameters(T tunable_, stream_t stream_) : tunable(tunable_), stream(stream_) {}
            T tunable;
            stream_t stream;
            };

        T checkParameter(const LaunchParameters& params) const
            {
            if (m_valid.count(params.tunable) == 0)
                throw tuning_error();
            return params.tunable;
            }

    private:
        std::set<T> m_valid;
    };

} // end namespace neighbor

#endif // NEIGHBOR_TUNABLE_H_
